// bvh_build.h -- host interface of the device acceleration-structure build.
#pragma once

#include <hip/hip_runtime.h>
#include <stdint.h>

#include <string>
#include <vector>

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
    float4* tris = nullptr;     // 3 float4 per triangle, leaf order, + 1 degenerate triangle at N (owned; gather only)
    uint32_t* order = nullptr;  // leaf order -> primitive (owned; without gather only)
    uint32_t root = kEmptyRoot;
    uint32_t num_nodes = 0, num_leaves = 0, max_depth = 0, max_leaf = 0;   // of the binary SAH tree
    uint32_t num_wide = 0, wide_depth = 0;                                    // of the 4-wide tree
    float root_lo[3] = {0.f, 0.f, 0.f}, root_hi[3] = {0.f, 0.f, 0.f};      // bounds of all primitives
    std::vector<uint32_t> level_begin;   // first wide node of each level (+ the total at the end)
    double sah = 0.0;
};

struct BvhOptions {
    uint32_t max_leaf = kMaxLeaf;   // primitives per leaf at most (1..kMaxLeaf)
    bool gather = true;             // write the leaf-ordered triangles (else return the leaf order)
};

// world-space triangles (3 float4 each: xyz | prim id, inst id, mesh prim) and
// per-primitive bounds, in instance-major order
hipError_t flatten_scene(const InstBuild* d_insts, uint32_t ninst, const MeshRef* d_meshes, uint32_t N,
                         float4* d_tri_world, float4* d_plo, float4* d_phi, hipStream_t s);

// binned SAH BVH over N flattened triangles; synchronous on s
// (without gather, d_tri_world is not read and may be null)
hipError_t build_bvh(const float4* d_tri_world, const float4* d_plo, const float4* d_phi, uint32_t N,
                     hipStream_t s, BuiltBVH* out, std::string* err, const BvhOptions& opt = BvhOptions());

// ---- two-level scenes (tlas.hip) ----
// adds node_base / tri_base to the refs of a bottom-level tree of n wide
// nodes that was copied to nodes[node_base ...] and its triangles to tri_base
hipError_t relocate_blas(float4* nodes, uint32_t n, uint32_t node_base, uint32_t tri_base, hipStream_t s);
// turns the leaf refs of a top-level tree (n wide nodes at nodes[0], one
// primitive per leaf, order = its leaf order over nprim primitives) into
// instance refs kInstLeaf | inst_of[primitive]; empty slots get null_inst
hipError_t tlas_leaves(float4* nodes, uint32_t n, const uint32_t* order, uint32_t nprim, const uint32_t* inst_of,
                       uint32_t null_inst, hipStream_t s);
// refits the top-level boxes from ibox (world lo, hi float4 per instance),
// deepest level first; null_ref = kInstLeaf | null instance
hipError_t refit_tlas(float4* nodes, const std::vector<uint32_t>& level_begin, const float4* ibox, uint32_t null_ref,
                      hipStream_t s);

}  // namespace rtx
