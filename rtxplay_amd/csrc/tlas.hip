// tlas.hip -- assembly and refit of a two-level scene: the instance
// acceleration structure (IAS) over per-mesh geometry structures (GAS) of
// optx/scene.cxx:38-294, here one 4-wide BVH per mesh over its object-space
// triangles (bottom level) and one over the instances' world boxes (top
// level), all in one node array and one triangle array.
//
// The trees come from build_bvh (bvh_build.hip).  Assembly copies each
// bottom-level tree into the scene arrays and adds the array offsets to its
// refs; the top level's leaves (one instance each) become instance refs.
// Scene::update (optx/scene.cxx:275-294, an OptiX refit) becomes a refit of
// the top level's boxes, level by level from the deepest, with the topology
// kept.
#include <hip/hip_runtime.h>

#include "bvh_build.h"

namespace rtx {

namespace {

__device__ __forceinline__ float comp(const float4& v, int j) {
    return j == 0 ? v.x : (j == 1 ? v.y : (j == 2 ? v.z : v.w));
}
__device__ __forceinline__ void set_comp(float4& v, int j, float x) {
    if (j == 0) v.x = x; else if (j == 1) v.y = x; else if (j == 2) v.z = x; else v.w = x;
}

// bottom-level tree copied to node index node_base / triangle index tri_base
__global__ void k_relocate(float4* __restrict__ nodes, uint32_t n, uint32_t node_base, uint32_t tri_base) {
    const uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    float4* o = nodes + (size_t)kWideNodeF4 * (node_base + i);
    float4 rf = o[6];
    for (int j = 0; j < 4; j++) {
        uint32_t r = __float_as_uint(comp(rf, j));
        if (r & kLeafFlag) r += tri_base;           // first triangle (bits 0..26; the host checks the range)
        else if (r != kNoChild) r += node_base;
        set_comp(rf, j, __uint_as_float(r));
    }
    o[6] = rf;
}

// top-level tree: leaf k (one primitive, order[k]) holds instance
// inst_of[order[k]]; the empty slots' leaf (first == nprim) the null instance
__global__ void k_tlas_leaves(float4* __restrict__ nodes, uint32_t n, const uint32_t* __restrict__ order,
                              uint32_t nprim, const uint32_t* __restrict__ inst_of, uint32_t null_inst) {
    const uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    float4* o = nodes + (size_t)kWideNodeF4 * i;
    float4 rf = o[6];
    for (int j = 0; j < 4; j++) {
        const uint32_t r = __float_as_uint(comp(rf, j));
        if (!(r & kLeafFlag)) continue;
        const uint32_t first = r & kLeafFirstMask;
        const uint32_t id = first < nprim ? inst_of[order[first]] : null_inst;
        set_comp(rf, j, __uint_as_float(kInstLeaf | id));
    }
    o[6] = rf;
}

// one level of the top-level refit: each child slot gets the instance's world
// box (ibox: lo, hi per instance) or the union of the inner child's slots
// (refit before, one level deeper); slots of the null instance keep its box
// and are left out of unions
__global__ void k_refit4(float4* __restrict__ nodes, uint32_t begin, uint32_t n, const float4* __restrict__ ibox,
                         uint32_t null_ref) {
    const uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    float4* o = nodes + (size_t)kWideNodeF4 * (begin + i);
    const float4 rf = o[6];
    float4 lx = rf, hx = rf, ly = rf, hy = rf, lz = rf, hz = rf;   // every component is set below
    for (int j = 0; j < 4; j++) {
        const uint32_t r = __float_as_uint(comp(rf, j));
        float bl[3], bh[3];
        if ((r & kInstLeaf) == kInstLeaf) {
            const uint32_t id = r & kLeafFirstMask;
            const float4 a = ibox[2 * (size_t)id], b = ibox[2 * (size_t)id + 1];
            bl[0] = a.x; bl[1] = a.y; bl[2] = a.z;
            bh[0] = b.x; bh[1] = b.y; bh[2] = b.z;
        } else {
            const float4* c = nodes + (size_t)kWideNodeF4 * r;
            const float4 cl[3] = {c[0], c[2], c[4]}, ch[3] = {c[1], c[3], c[5]};
            const float4 crf = c[6];
            for (int k = 0; k < 3; k++) { bl[k] = __builtin_huge_valf(); bh[k] = -__builtin_huge_valf(); }
            for (int q = 0; q < 4; q++) {
                if (__float_as_uint(comp(crf, q)) == null_ref) continue;
                for (int k = 0; k < 3; k++) {
                    bl[k] = fminf(bl[k], comp(cl[k], q));
                    bh[k] = fmaxf(bh[k], comp(ch[k], q));
                }
            }
        }
        set_comp(lx, j, bl[0]); set_comp(hx, j, bh[0]);
        set_comp(ly, j, bl[1]); set_comp(hy, j, bh[1]);
        set_comp(lz, j, bl[2]); set_comp(hz, j, bh[2]);
    }
    o[0] = lx; o[1] = hx; o[2] = ly; o[3] = hy; o[4] = lz; o[5] = hz;
}

inline uint32_t blocks256(uint32_t n) { return (n + 255u) / 256u; }

}  // namespace

hipError_t relocate_blas(float4* nodes, uint32_t n, uint32_t node_base, uint32_t tri_base, hipStream_t s) {
    if (n == 0) return hipSuccess;
    hipLaunchKernelGGL(k_relocate, dim3(blocks256(n)), dim3(256), 0, s, nodes, n, node_base, tri_base);
    return hipGetLastError();
}

hipError_t tlas_leaves(float4* nodes, uint32_t n, const uint32_t* order, uint32_t nprim, const uint32_t* inst_of,
                       uint32_t null_inst, hipStream_t s) {
    if (n == 0) return hipSuccess;
    hipLaunchKernelGGL(k_tlas_leaves, dim3(blocks256(n)), dim3(256), 0, s, nodes, n, order, nprim, inst_of,
                       null_inst);
    return hipGetLastError();
}

hipError_t refit_tlas(float4* nodes, const std::vector<uint32_t>& level_begin, const float4* ibox, uint32_t null_ref,
                      hipStream_t s) {
    // level_begin: first node of each level, then the node count
    for (size_t l = level_begin.size(); l-- > 1;) {
        const uint32_t b = level_begin[l - 1], n = level_begin[l] - b;
        if (n == 0) continue;
        hipLaunchKernelGGL(k_refit4, dim3(blocks256(n)), dim3(256), 0, s, nodes, b, n, ibox, null_ref);
        const hipError_t e = hipGetLastError();
        if (e != hipSuccess) return e;
    }
    return hipSuccess;
}

}  // namespace rtx
