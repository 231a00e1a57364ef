// kernels.h -- host-side launchers of the render, query and post-process kernels.
#pragma once

#include <hip/hip_runtime.h>
#include <stdint.h>

#include "rtx.h"
#include "rtx_internal.h"

namespace rtx {

constexpr int kRenderBlock = 256;
constexpr int kTraceBlock = 256;

// per-context scratch for the persistent kernels
struct LaunchScratch {
    uint32_t* counter = nullptr;   // work counter (device)
    uint32_t* gstack = nullptr;    // traversal stack overflow (device)
    uint32_t gstride = 0;          // lanes covered by gstack (grid * block)
    uint32_t grid = 0;             // persistent grid size in blocks
    uint32_t shade_batch = 40;     // render: finished lanes that trigger a wave's shading step
    uint32_t inner_frac = 36;      // traversal: node phase ends below inner_frac/64 of the going lanes
    uint32_t tile_w = 8;           // render: work tile of tile_w x (64 / tile_w) pixels
    uint32_t stack_lds = kStackLDS;   // render: traversal stack entries per lane in LDS (16 or 32)
    uint32_t grid_max = 0;         // blocks the overflow stack area was sized for
    uint32_t gstack_entries = 0;   // overflow entries per lane (beyond the smallest LDS part)
    uint32_t* tile_cost = nullptr;   // render: rays per tile of the one-sample probe (device)
    uint32_t* tile_order = nullptr;  // render: the tiles in descending probe cost (device)
    uint32_t tile_cap = 0;           // tiles the two arrays hold
    uint32_t iterative = 0;          // render: 1 = the iterative programs' semantics, 0 = the recursive ones
};

// Plain frames of at least kOrderMinSpp samples per pixel and at least
// kOrderMinPixelsPerLane pixels per lane of the persistent grid render their
// tiles longest first: a one-sample probe launch counts each tile's rays, and
// the frame takes its tiles in descending order of that count (stable, so
// equal counts keep their order).  A lane renders whole pixels, so at the end
// of a frame a wave waits for its most expensive last pixel; longest first
// makes the last pixels cheap ones.  A frame with fewer pixels than that has
// no such tail to shorten, and the probe would only add its time.  No pixel
// depends on the order.
constexpr uint32_t kOrderMinSpp = 16;
constexpr uint32_t kOrderMinPixelsPerLane = 2;
inline bool frame_ordered(const rtx_lp_general& lp, const LaunchScratch& ls) {
    return lp.spp >= kOrderMinSpp && (uint64_t)lp.image_w * lp.image_h >= (uint64_t)kOrderMinPixelsPerLane * ls.gstride;
}
// tiles of a frame (64 work slots each) for a tile width
inline uint32_t frame_tiles(const rtx_lp_general& lp, uint32_t tile_w) {
    const uint32_t th = 64u / tile_w;
    return ((lp.image_w + tile_w - 1u) / tile_w) * ((lp.image_h + th - 1u) / th);
}

// One frame of a launch: scene, launch parameters and its slice of the
// global work index space (64 slots per tile).
struct Job {
    DevScene sc;
    rtx_lp_general lp;
    float3 ambient;
    uint32_t tiles_x, tiles_y;
    uint32_t work_begin, work_end;
};

Job make_job(const DevScene& sc, const rtx_lp_general& lp, float3 ambient, uint32_t tile_w, uint32_t work_begin);

// Optional per-pixel traversal statistics (node visits, triangle tests).
struct TraceCounters {
    uint64_t* nodes = nullptr;
    uint64_t* tris = nullptr;
    uint64_t* wave_cycles = nullptr;   // [0] cycles in traversal, [1] cycles in the whole loop (sum over waves)
};

hipError_t launch_render(const DevScene& sc, const rtx_lp_general& lp, float3 ambient, const LaunchScratch& ls,
                         hipStream_t s);
hipError_t launch_render_counted(const DevScene& sc, const rtx_lp_general& lp, float3 ambient,
                                 const LaunchScratch& ls, const TraceCounters& tc, hipStream_t s);
// all jobs of a batch in one persistent launch (d_jobs sorted by work_begin)
hipError_t launch_render_batch(const Job* d_jobs, const Job& first, uint32_t njobs, uint32_t nwork,
                               const LaunchScratch& ls, uint32_t tile_w, hipStream_t s);
// adjoint of a frame: dparams (float4 per instance: d albedo rgb, d fuzz) =
// dL/d params given dL/d rawRGB (lp.rawRGB holds the forward image); tape:
// depth x kAdjointTapeWords words per lane (one record per shaded hit: at
// most depth - 1 recursive, depth iterative), lane stride tape_stride
constexpr uint32_t kAdjointTapeWords = 14;
constexpr uint32_t kAdjointLdsInsts = 2048;   // instances whose gradients a block sums in LDS first
hipError_t launch_adjoint(const DevScene& sc, const rtx_lp_general& lp, float3 ambient, const LaunchScratch& ls,
                          const rtx_float3* dldi, float* dparams, uint32_t* tape, uint32_t tape_stride, hipStream_t s);
hipError_t launch_set_materials(const DevScene& sc, const float* params, hipStream_t s);
hipError_t launch_get_materials(const DevScene& sc, float* params, hipStream_t s);
hipError_t launch_pick(const DevScene& sc, const rtx_lp_general& lp, const LaunchScratch& ls, hipStream_t s);
hipError_t launch_trace(const DevScene& sc, const rtx_ray* rays, uint32_t n, rtx_hit* hits,
                        const LaunchScratch& ls, hipStream_t s);
hipError_t launch_rng(const uint32_t* seeds, uint32_t n, uint32_t k, float* out, hipStream_t s);
hipError_t launch_pp(const rtx_float3* src, rtx_uchar4* dst, int w, int h, bool srgb, hipStream_t s);

// occupancy-derived persistent grid (blocks) for the render / trace kernels
uint32_t render_blocks_per_cu();
uint32_t trace_blocks_per_cu();

}  // namespace rtx
