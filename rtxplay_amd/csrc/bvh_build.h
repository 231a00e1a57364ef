// bvh_build.h -- host interface of the device acceleration-structure build.
#pragma once

#include <hip/hip_runtime.h>
#include <stdint.h>

#include <string>

#include "rtx_internal.h"

namespace rtx {

// one mesh as uploaded by rtx_mesh_add (device pointers)
struct MeshRef {
    const float* vces;     // 3 floats per vertex
    const uint32_t* ices;  // 3 indices per triangle
    uint32_t nv, ni;
};

// one instance for the flatten kernel: 3x4 row-major transform, mesh, first
// world-space triangle of the instance
struct InstBuild {
    float xf[12];
    uint32_t mesh;
    uint32_t tri_base;
    uint32_t pad[2];
};

struct BuiltBVH {
    float4* nodes = nullptr;    // kWideNodeF4 float4 per 4-wide node (owned)
    float4* tris = nullptr;     // 3 float4 per triangle, leaf order, + 1 degenerate triangle at N (owned)
    uint32_t root = kEmptyRoot;
    uint32_t num_nodes = 0, num_leaves = 0, max_depth = 0, max_leaf = 0;   // of the binary SAH tree
    uint32_t num_wide = 0, wide_depth = 0;                                    // of the 4-wide tree
    double sah = 0.0;
};

// world-space triangles (3 float4 each: xyz | prim id, inst id, mesh prim) and
// per-primitive bounds, in instance-major order
hipError_t flatten_scene(const InstBuild* d_insts, uint32_t ninst, const MeshRef* d_meshes, uint32_t N,
                         float4* d_tri_world, float4* d_plo, float4* d_phi, hipStream_t s);

// binned SAH BVH over N flattened triangles; synchronous on s
hipError_t build_bvh(const float4* d_tri_world, const float4* d_plo, const float4* d_phi, uint32_t N,
                     hipStream_t s, BuiltBVH* out, std::string* err);

}  // namespace rtx
