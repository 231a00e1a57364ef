// capi.cpp -- the extern "C" boundary (include/rtx.h): context, geometry,
// acceleration-structure build, launch, post-process and memory plumbing.
//
// Host-side counterparts: Scene (optx/scene.cxx), Launcher::ignite
// (optx/launcher.cxx:49-77), Camera::set (optx/camera.h:30-48), Sphere
// (optx/sphere.cxx) and pp_sRGB/pp_none (optx/postproc.cu:49-61).  Errors are
// status codes plus a per-thread message (the reference throws
// std::runtime_error from CUDA_CHECK/OPTX_CHECK, optx/util_cpu.h:18-53).
#include <hip/hip_runtime.h>

#include <algorithm>
#include <array>
#include <cmath>
#include <cstring>
#include <mutex>
#include <set>
#include <string>
#include <unordered_map>
#include <vector>

#include "bvh_build.h"
#include "host_util.h"
#include "kernels.h"
#include "rtx.h"
#include "rtx_internal.h"

namespace {

thread_local std::string g_err;

rtx_status fail(rtx_status st, const std::string& msg) {
    g_err = msg;
    return st;
}

}  // namespace

rtx_status rtx::set_error(rtx_status st, const std::string& msg) { return fail(st, msg); }

namespace {

rtx_status hip_fail(hipError_t e, const char* what) {
    return fail(e == hipErrorOutOfMemory ? RTX_ERROR_OUT_OF_MEMORY : RTX_ERROR_HIP,
                std::string(what) + ": " + hipGetErrorString(e));
}

#define HIP_OK(expr)                                         \
    do {                                                     \
        hipError_t e_ = (expr);                              \
        if (e_ != hipSuccess) return hip_fail(e_, #expr);    \
    } while (0)

struct MeshRec {
    float* d_v = nullptr;
    uint32_t* d_i = nullptr;
    uint32_t nv = 0, ni = 0;
};

struct ThingRec {
    rtx_thing thing;
    uint32_t mesh;
    float xf[12];
};

struct Built {
    rtx::DevScene hdr;                 // host copy of what the handle points at
    rtx::DevScene* d_hdr = nullptr;    // device copy; its address is the handle
    float4* nodes = nullptr;
    float4* tris = nullptr;
    float4* insts = nullptr;
    float4* tri_world = nullptr;       // flat: original-order world triangles (3 float4 each)
    int device = 0;
    uint32_t stack_need = 1;           // traversal stack entries per lane this scene can need
    rtx_scene_stats stats{};
    // two-level scenes
    bool two_level = false;
    float4* xinst = nullptr;           // instance records (kInstF4 float4 each, + the null instance)
    float4* ibox = nullptr;            // instance world boxes (lo, hi), + the null instance
    std::vector<uint32_t> tlas_levels; // first top-level node of each level, then the count
    std::vector<uint32_t> blas_root;   // per mesh: bottom-level root in the scene arrays
    std::vector<float> blas_box;       // per mesh: object-space bounds lo.xyz, hi.xyz
    uint32_t null_root = 0;            // a one-triangle leaf of a degenerate triangle
    std::vector<rtx::InstBuild> ib;    // transforms and first global triangle per instance at the last build/update
    void release() {
        if (d_hdr) (void)hipFree(d_hdr);
        if (nodes) (void)hipFree(nodes);
        if (tris) (void)hipFree(tris);
        if (insts) (void)hipFree(insts);
        if (tri_world) (void)hipFree(tri_world);
        if (xinst) (void)hipFree(xinst);
        if (ibox) (void)hipFree(ibox);
        d_hdr = nullptr; nodes = tris = insts = tri_world = xinst = ibox = nullptr;
    }
};

// process-wide registry of live traversable handles (device address -> scene)
std::mutex g_reg_mu;
std::unordered_map<uint64_t, Built*> g_registry;

Built* lookup(uint64_t handle) {
    std::lock_guard<std::mutex> lk(g_reg_mu);
    auto it = g_registry.find(handle);
    return it == g_registry.end() ? nullptr : it->second;
}

const float kIdentity[12] = {1.f, 0.f, 0.f, 0.f, 0.f, 1.f, 0.f, 0.f, 0.f, 0.f, 1.f, 0.f};

// device buffer freed at scope exit (build temporaries)
template <typename T>
struct Tmp {
    T* p = nullptr;
    ~Tmp() { if (p) (void)hipFree(p); }
};

}  // namespace

// ---- scratch of rtx_launch_noalloc ------------------------------------------
// One static device array holds every scratch buffer of a no-allocation
// launch: the work counter (first 256 bytes), the probe's rays per tile and
// the tile order (kNoAllocTiles entries each), then the traversal stack's
// overflow area (the rest).  Each GPU the library runs on holds its own copy
// until the process ends.  Its device address is looked up once per device
// (noalloc_setup, from rtx_ctx_create with that device current); host code
// uses that address, never the array's name.
constexpr size_t kNoAllocBytes = (size_t)48 << 20;
constexpr uint32_t kNoAllocTiles = 1u << 18;   // frames of up to 2^24 pixels render in probe order
constexpr size_t kNoAllocStackOffset = 256 + 2 * sizeof(uint32_t) * (size_t)kNoAllocTiles;
static __device__ __align__(256) uint8_t scratch[kNoAllocBytes];

namespace {
constexpr int kNoAllocMaxDevices = 64;
std::mutex g_scratch_mu;
bool g_scratch_done[kNoAllocMaxDevices] = {};
uint8_t* g_scratch_ptr[kNoAllocMaxDevices] = {};
hipError_t g_scratch_err[kNoAllocMaxDevices] = {};

// with `device` current: its copy's address, once per device
void noalloc_setup(int device) {
    if (device < 0 || device >= kNoAllocMaxDevices) return;
    std::lock_guard<std::mutex> lk(g_scratch_mu);
    if (g_scratch_done[device]) return;
    void* p = nullptr;
    const hipError_t e = hipGetSymbolAddress(&p, HIP_SYMBOL(scratch));
    g_scratch_err[device] = e;
    g_scratch_ptr[device] = e == hipSuccess ? (uint8_t*)p : nullptr;
    g_scratch_done[device] = true;
}
}  // namespace

struct rtx_ctx {
    int device = 0;
    int num_cus = 0;
    std::vector<MeshRec> meshes;
    std::vector<ThingRec> things;
    Built* built = nullptr;
    rtx_float3 ambient{.5f, .7f, 1.f};
    rtx::LaunchScratch render_ls, trace_ls;
    int build_mode = RTX_BUILD_FLAT;
    uint32_t* tape = nullptr;       // adjoint path records
    uint32_t tape_depth = 0;        // depth the tape was sized for
};

extern "C" {

const char* rtx_last_error(void) { return g_err.c_str(); }

static rtx_status alloc_scratch(rtx::LaunchScratch& ls, uint32_t blocks_per_cu, int num_cus, int block) {
    ls.grid = blocks_per_cu * (uint32_t)num_cus;
    ls.grid_max = ls.grid;
    ls.gstride = ls.grid * (uint32_t)block;
    HIP_OK(hipMalloc((void**)&ls.counter, sizeof(uint32_t)));
    return RTX_SUCCESS;
}

// the traversal stack overflow area (tens of MB) is allocated by the first
// launch that needs it, so contexts that only hold a scene for a batch cost
// nothing beyond the scene; it grows when a scene needs a deeper stack
static rtx_status ensure_gstack(rtx::LaunchScratch& ls, uint32_t need) {
    const uint32_t over = need > (uint32_t)rtx::kMinStackLDS ? need - (uint32_t)rtx::kMinStackLDS : 1u;
    if (ls.gstack && ls.gstack_entries >= over) return RTX_SUCCESS;
    if (ls.gstack) {
        HIP_OK(hipDeviceSynchronize());
        HIP_OK(hipFree(ls.gstack));
        ls.gstack = nullptr;
        ls.gstack_entries = 0;
    }
    HIP_OK(hipMalloc((void**)&ls.gstack, sizeof(uint32_t) * (size_t)over * ls.gstride));
    ls.gstack_entries = over;
    return RTX_SUCCESS;
}

// the probe cost and tile order arrays of ordered frames (a few hundred KB),
// allocated by the first launch that needs them and grown with the frame
static rtx_status ensure_order(rtx::LaunchScratch& ls, uint32_t ntiles) {
    if (ls.tile_cost && ls.tile_cap >= ntiles) return RTX_SUCCESS;
    if (ls.tile_cost) {
        HIP_OK(hipDeviceSynchronize());
        HIP_OK(hipFree(ls.tile_cost));
        HIP_OK(hipFree(ls.tile_order));
        ls.tile_cost = ls.tile_order = nullptr;
        ls.tile_cap = 0;
    }
    HIP_OK(hipMalloc((void**)&ls.tile_cost, sizeof(uint32_t) * (size_t)ntiles));
    hipError_t e = hipMalloc((void**)&ls.tile_order, sizeof(uint32_t) * (size_t)ntiles);
    if (e != hipSuccess) {
        (void)hipFree(ls.tile_cost);
        ls.tile_cost = nullptr;
        return hip_fail(e, "hipMalloc(tile_order)");
    }
    ls.tile_cap = ntiles;
    return RTX_SUCCESS;
}

rtx_status rtx_ctx_create(int device, rtx_ctx** out) {
    if (!out) return fail(RTX_ERROR_INVALID_VALUE, "rtx_ctx_create: ctx is NULL");
    int n = 0;
    HIP_OK(hipGetDeviceCount(&n));
    if (device < 0 || device >= n) return fail(RTX_ERROR_INVALID_VALUE, "rtx_ctx_create: no such device");
    HIP_OK(hipSetDevice(device));
    hipDeviceProp_t prop;
    HIP_OK(hipGetDeviceProperties(&prop, device));
    noalloc_setup(device);
    rtx_ctx* c = new rtx_ctx();
    c->device = device;
    c->num_cus = prop.multiProcessorCount;
    rtx_status st = alloc_scratch(c->render_ls, rtx::render_blocks_per_cu(), c->num_cus, rtx::kRenderBlock);
    if (st == RTX_SUCCESS)
        st = alloc_scratch(c->trace_ls, rtx::trace_blocks_per_cu(), c->num_cus, rtx::kTraceBlock);
    if (st != RTX_SUCCESS) {
        rtx_ctx_destroy(c);
        return st;
    }
    *out = c;
    return RTX_SUCCESS;
}

static void drop_built(rtx_ctx* c) {
    if (!c->built) return;
    {
        std::lock_guard<std::mutex> lk(g_reg_mu);
        g_registry.erase((uint64_t)(uintptr_t)c->built->d_hdr);
    }
    c->built->release();
    delete c->built;
    c->built = nullptr;
}

rtx_status rtx_ctx_destroy(rtx_ctx* c) {
    if (!c) return RTX_SUCCESS;
    (void)hipSetDevice(c->device);
    (void)hipDeviceSynchronize();
    drop_built(c);
    for (auto& m : c->meshes) {
        if (m.d_v) (void)hipFree(m.d_v);
        if (m.d_i) (void)hipFree(m.d_i);
    }
    for (auto* ls : {&c->render_ls, &c->trace_ls}) {
        if (ls->counter) (void)hipFree(ls->counter);
        if (ls->gstack) (void)hipFree(ls->gstack);
        if (ls->tile_cost) (void)hipFree(ls->tile_cost);
        if (ls->tile_order) (void)hipFree(ls->tile_order);
    }
    if (c->tape) (void)hipFree(c->tape);
    delete c;
    return RTX_SUCCESS;
}

// ---- camera: optx/camera.h:30-48 ------------------------------------------
rtx_status rtx_camera_set(rtx_camera* c, rtx_float3 eye, rtx_float3 pat, rtx_float3 vup, float fov, float aspratio,
                          float aperture, float fostance) {
    if (!c) return fail(RTX_ERROR_INVALID_VALUE, "rtx_camera_set: cam is NULL");
    auto sub = [](rtx_float3 a, rtx_float3 b) { return rtx_float3{a.x - b.x, a.y - b.y, a.z - b.z}; };
    auto scl = [](float s, rtx_float3 a) { return rtx_float3{s * a.x, s * a.y, s * a.z}; };
    auto dot = [](rtx_float3 a, rtx_float3 b) { return a.x * b.x + a.y * b.y + a.z * b.z; };
    auto cross = [](rtx_float3 a, rtx_float3 b) {
        return rtx_float3{a.y * b.z - a.z * b.y, a.z * b.x - a.x * b.z, a.x * b.y - a.y * b.x};
    };
    auto unit = [&](rtx_float3 a) { return scl(1.f / std::sqrt(dot(a, a)), a); };
    c->eye = eye; c->pat = pat; c->vup = vup;
    c->fov = fov; c->aspratio = aspratio; c->aperture = aperture; c->fostance = fostance;
    c->w = unit(sub(eye, pat));
    c->u = unit(cross(vup, c->w));
    c->v = cross(c->w, c->u);
    const float kPi = 3.14159265358979323846f;
    const float rad = fov * kPi / 180.f;
    const float h = 2.f * std::tan(.5f * rad);
    const float w = h * aspratio;
    c->hvec = scl(fostance * h / 2.f, c->v);
    c->wvec = scl(fostance * w / 2.f, c->u);
    c->dvec = scl(fostance, c->w);
    return RTX_SUCCESS;
}

// ---- sphere tessellation: optx/sphere.cxx:28-106 --------------------------
namespace {
struct V3 { float x, y, z; };
inline V3 operator+(V3 a, V3 b) { return {a.x + b.x, a.y + b.y, a.z + b.z}; }
inline V3 operator*(float s, V3 a) { return {s * a.x, s * a.y, s * a.z}; }
inline V3 unitv(V3 a) { return (1.f / std::sqrt(a.x * a.x + a.y * a.y + a.z * a.z)) * a; }

// recursive 4-way split of a spherical triangle; leaves emit 3 scaled vertices
void subdivide(std::vector<V3>& out, float radius, V3 a, V3 b, V3 c, uint32_t level) {
    if (level == 0) {
        out.push_back(radius * a);
        out.push_back(radius * b);
        out.push_back(radius * c);
        return;
    }
    const V3 ab = unitv(.5f * (a + b)), bc = unitv(.5f * (b + c)), ca = unitv(.5f * (c + a));
    subdivide(out, radius, a, ab, ca, level - 1);
    subdivide(out, radius, ab, b, bc, level - 1);
    subdivide(out, radius, ca, bc, c, level - 1);
    subdivide(out, radius, ab, bc, ca, level - 1);
}

// exact-value key (+0 and -0 are one vertex, as the reference's comparator has it)
struct KeyHash {
    size_t operator()(const std::array<uint32_t, 3>& k) const {
        return (size_t)k[0] * 0x9E3779B97F4A7C15ull ^ (size_t)k[1] * 0xC2B2AE3D27D4EB4Full ^ (size_t)k[2] * 0x165667B19E3779F9ull;
    }
};
inline uint32_t key_bits(float f) {
    if (f == 0.f) f = 0.f;
    uint32_t u;
    std::memcpy(&u, &f, 4);
    return u;
}
}  // namespace

rtx_status rtx_sphere_mesh(float radius, uint32_t ndiv, rtx_float3* vces, uint32_t vcap, uint32_t* nv,
                           rtx_uint3* ices, uint32_t icap, uint32_t* ni) {
    if (!nv || !ni) return fail(RTX_ERROR_INVALID_VALUE, "rtx_sphere_mesh: nv/ni is NULL");
    if (ndiv > 11) return fail(RTX_ERROR_LIMIT, "rtx_sphere_mesh: ndiv > 11");
    const uint64_t faces = 4ull << (2 * ndiv);
    const uint64_t verts = 2ull * (1ull << (2 * ndiv)) + 2ull;
    *nv = (uint32_t)verts;
    *ni = (uint32_t)faces;
    if (!vces && !ices) return RTX_SUCCESS;
    if (!vces || !ices || vcap < verts || icap < faces)
        return fail(RTX_ERROR_INVALID_VALUE, "rtx_sphere_mesh: output capacity too small");
    const float m = .57735026919f;   // optx/sphere.cxx:34
    const V3 v0{m, m, m}, v1{m, -m, -m}, v2{-m, -m, m}, v3{-m, m, -m};
    std::vector<V3> soup;
    soup.reserve(3 * faces);
    subdivide(soup, radius, v0, v1, v2, ndiv);
    subdivide(soup, radius, v0, v2, v3, ndiv);
    subdivide(soup, radius, v0, v3, v1, ndiv);
    subdivide(soup, radius, v3, v2, v1, ndiv);
    std::unordered_map<std::array<uint32_t, 3>, uint32_t, KeyHash> seen;
    seen.reserve(2 * verts);
    uint32_t count = 0;
    std::vector<uint32_t> idx(soup.size());
    for (size_t i = 0; i < soup.size(); i++) {
        const V3 p = soup[i];
        const std::array<uint32_t, 3> k{key_bits(p.x), key_bits(p.y), key_bits(p.z)};
        auto it = seen.find(k);
        if (it == seen.end()) {
            if (count >= verts) return fail(RTX_ERROR_LIMIT, "rtx_sphere_mesh: vertex count mismatch");
            seen.emplace(k, count);
            vces[count] = rtx_float3{p.x, p.y, p.z};
            idx[i] = count++;
        } else {
            idx[i] = it->second;
        }
    }
    for (uint64_t f = 0; f < faces; f++) ices[f] = rtx_uint3{idx[3 * f], idx[3 * f + 1], idx[3 * f + 2]};
    *nv = count;
    return RTX_SUCCESS;
}

// ---- geometry / instances (Scene) ------------------------------------------
rtx_status rtx_mesh_add(rtx_ctx* c, const rtx_float3* vces, uint32_t nv, const rtx_uint3* ices, uint32_t ni,
                        uint32_t* mesh_id) {
    if (!c || !mesh_id) return fail(RTX_ERROR_INVALID_VALUE, "rtx_mesh_add: NULL argument");
    if ((nv && !vces) || (ni && !ices)) return fail(RTX_ERROR_INVALID_VALUE, "rtx_mesh_add: NULL buffer");
    for (uint32_t i = 0; i < ni; i++)
        if (ices[i].x >= nv || ices[i].y >= nv || ices[i].z >= nv)
            return fail(RTX_ERROR_INVALID_VALUE, "rtx_mesh_add: index out of bounds at triangle " + std::to_string(i));
    HIP_OK(hipSetDevice(c->device));
    MeshRec m;
    m.nv = nv;
    m.ni = ni;
    HIP_OK(hipMalloc((void**)&m.d_v, sizeof(rtx_float3) * (nv ? nv : 1)));
    HIP_OK(hipMalloc((void**)&m.d_i, sizeof(rtx_uint3) * (ni ? ni : 1)));
    if (nv) HIP_OK(hipMemcpy(m.d_v, vces, sizeof(rtx_float3) * nv, hipMemcpyHostToDevice));
    if (ni) HIP_OK(hipMemcpy(m.d_i, ices, sizeof(rtx_uint3) * ni, hipMemcpyHostToDevice));
    *mesh_id = (uint32_t)c->meshes.size();
    c->meshes.push_back(m);
    return RTX_SUCCESS;
}

rtx_status rtx_thing_add(rtx_ctx* c, rtx_thing* thing, uint32_t mesh_id, uint32_t* thing_id) {
    if (!c || !thing || !thing_id) return fail(RTX_ERROR_INVALID_VALUE, "rtx_thing_add: NULL argument");
    if (mesh_id >= c->meshes.size()) return fail(RTX_ERROR_INVALID_HANDLE, "rtx_thing_add: unknown mesh id");
    if (thing->optics.type < 0 || thing->optics.type >= RTX_OPTICS_NUM)
        return fail(RTX_ERROR_INVALID_VALUE, "rtx_thing_add: unknown optics type");
    thing->vces = (rtx_float3*)c->meshes[mesh_id].d_v;    // optx/scene.cxx:191-192
    thing->ices = (rtx_uint3*)c->meshes[mesh_id].d_i;
    ThingRec r;
    r.thing = *thing;
    r.mesh = mesh_id;
    std::memcpy(r.xf, kIdentity, sizeof(r.xf));
    *thing_id = (uint32_t)c->things.size();
    c->things.push_back(r);
    return RTX_SUCCESS;
}

rtx_status rtx_thing_set(rtx_ctx* c, uint32_t id, const float transform[12]) {
    if (!c || !transform) return fail(RTX_ERROR_INVALID_VALUE, "rtx_thing_set: NULL argument");
    if (id >= c->things.size()) return fail(RTX_ERROR_INVALID_HANDLE, "rtx_thing_set: id out of range");
    std::memcpy(c->things[id].xf, transform, sizeof(float) * 12);
    return RTX_SUCCESS;
}

rtx_status rtx_thing_get(rtx_ctx* c, uint32_t id, float transform[12]) {
    if (!c || !transform) return fail(RTX_ERROR_INVALID_VALUE, "rtx_thing_get: NULL argument");
    if (id >= c->things.size()) return fail(RTX_ERROR_INVALID_HANDLE, "rtx_thing_get: id out of range");
    std::memcpy(transform, c->things[id].xf, sizeof(float) * 12);
    return RTX_SUCCESS;
}

rtx_status rtx_thing_at(rtx_ctx* c, uint32_t id, rtx_thing* thing) {
    if (!c || !thing) return fail(RTX_ERROR_INVALID_VALUE, "rtx_thing_at: NULL argument");
    if (id >= c->things.size()) return fail(RTX_ERROR_INVALID_HANDLE, "rtx_thing_at: id out of range");
    *thing = c->things[id].thing;
    return RTX_SUCCESS;
}

rtx_status rtx_thing_count(rtx_ctx* c, uint32_t* n) {
    if (!c || !n) return fail(RTX_ERROR_INVALID_VALUE, "rtx_thing_count: NULL argument");
    *n = (uint32_t)c->things.size();
    return RTX_SUCCESS;
}

rtx_status rtx_tuning_get(rtx_ctx* c, rtx_tuning* t) {
    if (!c || !t) return fail(RTX_ERROR_INVALID_VALUE, "rtx_tuning_get: NULL argument");
    std::memset(t, 0, sizeof(*t));
    t->shade_batch = c->render_ls.shade_batch;
    t->inner_frac = c->render_ls.inner_frac;
    t->tile_w = c->render_ls.tile_w;
    t->stack_lds = c->render_ls.stack_lds;
    t->blocks_per_cu = c->render_ls.grid / (uint32_t)c->num_cus;
    t->max_blocks_per_cu = c->render_ls.grid_max / (uint32_t)c->num_cus;
    return RTX_SUCCESS;
}

rtx_status rtx_tuning_set(rtx_ctx* c, const rtx_tuning* t) {
    if (!c || !t) return fail(RTX_ERROR_INVALID_VALUE, "rtx_tuning_set: NULL argument");
    if (t->shade_batch > 64) return fail(RTX_ERROR_INVALID_VALUE, "rtx_tuning_set: shade_batch > 64");
    if (t->inner_frac > 64) return fail(RTX_ERROR_INVALID_VALUE, "rtx_tuning_set: inner_frac > 64");
    if (t->tile_w > 64 || (t->tile_w & (t->tile_w - 1)))
        return fail(RTX_ERROR_INVALID_VALUE, "rtx_tuning_set: tile_w must be a power of two <= 64");
    if (t->stack_lds && t->stack_lds != 16 && t->stack_lds != 32)
        return fail(RTX_ERROR_INVALID_VALUE, "rtx_tuning_set: stack_lds must be 16 or 32");
    if (t->blocks_per_cu * (uint32_t)c->num_cus > c->render_ls.grid_max)
        return fail(RTX_ERROR_INVALID_VALUE, "rtx_tuning_set: blocks_per_cu above the occupancy limit " +
                                                 std::to_string(c->render_ls.grid_max / (uint32_t)c->num_cus));
    if (t->tile_w) c->render_ls.tile_w = t->tile_w;
    if (t->stack_lds) c->render_ls.stack_lds = t->stack_lds;
    if (t->blocks_per_cu) c->render_ls.grid = t->blocks_per_cu * (uint32_t)c->num_cus;
    if (t->shade_batch) c->render_ls.shade_batch = t->shade_batch;
    if (t->inner_frac) c->render_ls.inner_frac = t->inner_frac;
    return RTX_SUCCESS;
}

rtx_status rtx_set_miss_color(rtx_ctx* c, rtx_float3 ambient) {
    if (!c) return fail(RTX_ERROR_INVALID_VALUE, "rtx_set_miss_color: ctx is NULL");
    c->ambient = ambient;
    return RTX_SUCCESS;
}

rtx_status rtx_set_semantics(rtx_ctx* c, int mode) {
    if (!c) return fail(RTX_ERROR_INVALID_VALUE, "rtx_set_semantics: ctx is NULL");
    if (mode != RTX_SEMANTICS_RECURSIVE && mode != RTX_SEMANTICS_ITERATIVE)
        return fail(RTX_ERROR_INVALID_VALUE, "rtx_set_semantics: unknown mode");
    c->render_ls.iterative = mode == RTX_SEMANTICS_ITERATIVE ? 1u : 0u;
    return RTX_SUCCESS;
}

rtx_status rtx_get_semantics(rtx_ctx* c, int* mode) {
    if (!c || !mode) return fail(RTX_ERROR_INVALID_VALUE, "rtx_get_semantics: NULL argument");
    *mode = c->render_ls.iterative ? RTX_SEMANTICS_ITERATIVE : RTX_SEMANTICS_RECURSIVE;
    return RTX_SUCCESS;
}

// Per instance: the flatten record (transform, mesh, first global triangle)
// and the material record of the scene (2 float4); N = triangles of all
// instances.
static void instance_records(const rtx_ctx* c, std::vector<rtx::InstBuild>& ib, std::vector<float4>& irec,
                             uint64_t& N) {
    const uint32_t ninst = (uint32_t)c->things.size();
    ib.assign(ninst ? ninst : 1, rtx::InstBuild{});
    irec.assign(2 * (size_t)(ninst ? ninst : 1), make_float4(0.f, 0.f, 0.f, 0.f));
    N = 0;
    for (uint32_t i = 0; i < ninst; i++) {
        const ThingRec& t = c->things[i];
        std::memcpy(ib[i].xf, t.xf, sizeof(t.xf));
        ib[i].mesh = t.mesh;
        ib[i].tri_base = (uint32_t)N;
        ib[i].pad[0] = ib[i].pad[1] = 0;
        N += c->meshes[t.mesh].ni;
        const rtx_optics& op = t.thing.optics;
        float e = 0.f;
        rtx_float3 alb{0.f, 0.f, 0.f};
        if (op.type == RTX_OPTICS_DIFFUSE) alb = op.diffuse.albedo;
        else if (op.type == RTX_OPTICS_REFLECT) { alb = op.reflect.albedo; e = op.reflect.fuzz; }
        else e = op.refract.index;
        int32_t ty = op.type;
        float tyf;
        std::memcpy(&tyf, &ty, 4);
        irec[2 * i] = make_float4(alb.x, alb.y, alb.z, e);
        irec[2 * i + 1] = make_float4(t.xf[3], t.xf[7], t.xf[11], tyf);   // center: optx/optics_r.cu:376-380
    }
}

// Flattens every instance and builds the BVH into a new Built (not registered).
static rtx_status build_scene_flat(rtx_ctx* c, hipStream_t s, Built** out) {
    HIP_OK(hipSetDevice(c->device));
    const uint32_t ninst = (uint32_t)c->things.size();
    std::vector<rtx::InstBuild> ib;
    std::vector<float4> irec;
    uint64_t N = 0;
    instance_records(c, ib, irec, N);
    if (N > rtx::kMaxPrims) return fail(RTX_ERROR_LIMIT, "rtx_scene_build: more than 2^27-1 triangles");

    hipEvent_t e0, e1;
    HIP_OK(hipEventCreate(&e0));
    HIP_OK(hipEventCreate(&e1));
    HIP_OK(hipEventRecord(e0, s));

    Built* b = new Built();
    b->device = c->device;
    auto bail = [&](rtx_status st) { b->release(); delete b; (void)hipEventDestroy(e0); (void)hipEventDestroy(e1); return st; };
    std::vector<rtx::MeshRef> mr(c->meshes.size() ? c->meshes.size() : 1);
    for (size_t m = 0; m < c->meshes.size(); m++)
        mr[m] = rtx::MeshRef{c->meshes[m].d_v, c->meshes[m].d_i, c->meshes[m].nv, c->meshes[m].ni};
    Tmp<rtx::InstBuild> d_ib;
    Tmp<rtx::MeshRef> d_mr;
    Tmp<float4> d_plo, d_phi;
    hipError_t e;
    if ((e = hipMalloc((void**)&d_ib.p, sizeof(rtx::InstBuild) * ib.size())) != hipSuccess) return bail(hip_fail(e, "hipMalloc(insts)"));
    if ((e = hipMalloc((void**)&d_mr.p, sizeof(rtx::MeshRef) * mr.size())) != hipSuccess) return bail(hip_fail(e, "hipMalloc(meshes)"));
    if ((e = hipMalloc((void**)&b->insts, sizeof(float4) * irec.size())) != hipSuccess) return bail(hip_fail(e, "hipMalloc(inst records)"));
    if ((e = hipMalloc((void**)&b->tri_world, sizeof(float4) * 3 * (N ? N : 1))) != hipSuccess) return bail(hip_fail(e, "hipMalloc(triangles)"));
    if ((e = hipMalloc((void**)&d_plo.p, sizeof(float4) * (N ? N : 1))) != hipSuccess) return bail(hip_fail(e, "hipMalloc(bounds)"));
    if ((e = hipMalloc((void**)&d_phi.p, sizeof(float4) * (N ? N : 1))) != hipSuccess) return bail(hip_fail(e, "hipMalloc(bounds)"));
    if ((e = hipMemcpyAsync(d_ib.p, ib.data(), sizeof(rtx::InstBuild) * ib.size(), hipMemcpyHostToDevice, s)) != hipSuccess) return bail(hip_fail(e, "upload insts"));
    if ((e = hipMemcpyAsync(d_mr.p, mr.data(), sizeof(rtx::MeshRef) * mr.size(), hipMemcpyHostToDevice, s)) != hipSuccess) return bail(hip_fail(e, "upload meshes"));
    if ((e = hipMemcpyAsync(b->insts, irec.data(), sizeof(float4) * irec.size(), hipMemcpyHostToDevice, s)) != hipSuccess) return bail(hip_fail(e, "upload inst records"));
    if ((e = rtx::flatten_scene(d_ib.p, ninst, d_mr.p, (uint32_t)N, b->tri_world, d_plo.p, d_phi.p, s)) != hipSuccess) return bail(hip_fail(e, "flatten"));
    rtx::BuiltBVH bvh;
    std::string err;
    if ((e = rtx::build_bvh(b->tri_world, d_plo.p, d_phi.p, (uint32_t)N, s, &bvh, &err)) != hipSuccess)
        return bail(fail(e == hipErrorOutOfMemory ? RTX_ERROR_OUT_OF_MEMORY : RTX_ERROR_HIP, "build_bvh: " + err));
    b->nodes = bvh.nodes;
    b->tris = bvh.tris;
    b->hdr.nodes = bvh.nodes;
    b->hdr.tris = bvh.tris;
    b->hdr.insts = b->insts;
    b->hdr.xinst = nullptr;
    b->hdr.root = bvh.root;
    b->hdr.num_prims = (uint32_t)N;
    b->hdr.num_nodes = bvh.num_wide;
    b->hdr.num_insts = ninst;
    if ((e = hipMalloc((void**)&b->d_hdr, sizeof(rtx::DevScene))) != hipSuccess) return bail(hip_fail(e, "hipMalloc(handle)"));
    if ((e = hipMemcpyAsync(b->d_hdr, &b->hdr, sizeof(rtx::DevScene), hipMemcpyHostToDevice, s)) != hipSuccess) return bail(hip_fail(e, "upload handle"));
    if ((e = hipEventRecord(e1, s)) != hipSuccess) return bail(hip_fail(e, "hipEventRecord"));
    if ((e = hipStreamSynchronize(s)) != hipSuccess) return bail(hip_fail(e, "hipStreamSynchronize"));
    float ms = 0.f;
    (void)hipEventElapsedTime(&ms, e0, e1);
    (void)hipEventDestroy(e0);
    (void)hipEventDestroy(e1);
    if ((uint64_t)bvh.num_wide * 16u * rtx::kWideNodeF4 >= (1ull << 32) || ((uint64_t)N + 1) * 48u >= (1ull << 32)) {
        b->release();
        delete b;
        return fail(RTX_ERROR_LIMIT, "rtx_scene_build: node or triangle array above 4 GiB (too many triangles)");
    }
    if (bvh.max_depth > (uint32_t)rtx::kMaxTreeDepth) {
        b->release();
        delete b;
        return fail(RTX_ERROR_LIMIT, "rtx_scene_build: tree deeper than the traversal stack");
    }
    b->stats.num_prims = (uint32_t)N;
    b->stats.num_nodes = bvh.num_nodes;
    b->stats.num_leaves = bvh.num_leaves;
    b->stats.max_depth = bvh.max_depth;
    b->stats.max_leaf = bvh.max_leaf;
    b->stats.sah_cost = (float)bvh.sah;
    b->stats.build_ms = ms;
    b->stats.device_bytes = sizeof(float4) * ((uint64_t)rtx::kWideNodeF4 * bvh.num_wide + 3ull * N + 2ull * ninst);
    b->stats.wide_nodes = bvh.num_wide;
    b->stats.wide_depth = bvh.wide_depth;
    b->stats.unique_prims = (uint32_t)N;
    b->ib = ib;
    // a step pushes at most 3 entries per level of the 4-wide tree
    b->stack_need = 3u * (bvh.wide_depth > 0 ? bvh.wide_depth : 1u);
    *out = b;
    return RTX_SUCCESS;
}

// ---- two-level scenes -------------------------------------------------------
namespace {

// owned results of build_bvh, freed at scope exit
struct BvhSet {
    std::vector<rtx::BuiltBVH> v;
    ~BvhSet() {
        for (auto& x : v) {
            if (x.nodes) (void)hipFree(x.nodes);
            if (x.tris) (void)hipFree(x.tris);
            if (x.order) (void)hipFree(x.order);
        }
    }
};

// One instance of a two-level scene: its record (world-to-object rows,
// object-to-world rows, (root, first global triangle, mesh, 0)) and its world
// box from the 8 corners of the mesh's object-space box, widened by 2^-20 of
// its largest coordinate so that the float rounding of the corners cannot cut
// off a hit.  A transform without an inverse makes the instance unhittable:
// the null root and a point box at its translation.
void tl_instance(const rtx::InstBuild& in, uint32_t root, uint32_t null_root, const float* box, float4* rec,
                 float4* wb) {
    double m[3][3], t[3];
    for (int r = 0; r < 3; r++) {
        for (int k = 0; k < 3; k++) m[r][k] = in.xf[4 * r + k];
        t[r] = in.xf[4 * r + 3];
    }
    const double c00 = m[1][1] * m[2][2] - m[1][2] * m[2][1];
    const double c01 = m[1][2] * m[2][0] - m[1][0] * m[2][2];
    const double c02 = m[1][0] * m[2][1] - m[1][1] * m[2][0];
    const double det = m[0][0] * c00 + m[0][1] * c01 + m[0][2] * c02;
    double inv[3][3] = {{0, 0, 0}, {0, 0, 0}, {0, 0, 0}};
    bool ok = std::isfinite(det) && det != 0.0;
    if (ok) {
        const double id = 1.0 / det;
        inv[0][0] = c00 * id;
        inv[0][1] = (m[0][2] * m[2][1] - m[0][1] * m[2][2]) * id;
        inv[0][2] = (m[0][1] * m[1][2] - m[0][2] * m[1][1]) * id;
        inv[1][0] = c01 * id;
        inv[1][1] = (m[0][0] * m[2][2] - m[0][2] * m[2][0]) * id;
        inv[1][2] = (m[0][2] * m[1][0] - m[0][0] * m[1][2]) * id;
        inv[2][0] = c02 * id;
        inv[2][1] = (m[0][1] * m[2][0] - m[0][0] * m[2][1]) * id;
        inv[2][2] = (m[0][0] * m[1][1] - m[0][1] * m[1][0]) * id;
    }
    float w2o[12];
    for (int r = 0; r < 3 && ok; r++) {
        double tr = 0.0;
        for (int k = 0; k < 3; k++) {
            w2o[4 * r + k] = (float)inv[r][k];
            tr -= inv[r][k] * t[k];
        }
        w2o[4 * r + 3] = (float)tr;
        for (int k = 0; k < 4; k++) ok = ok && std::isfinite(w2o[4 * r + k]);
    }
    if (!ok) std::memcpy(w2o, kIdentity, sizeof(w2o));
    for (int r = 0; r < 3; r++) {
        rec[r] = make_float4(w2o[4 * r], w2o[4 * r + 1], w2o[4 * r + 2], w2o[4 * r + 3]);
        rec[3 + r] = make_float4(in.xf[4 * r], in.xf[4 * r + 1], in.xf[4 * r + 2], in.xf[4 * r + 3]);
    }
    uint32_t meta[4] = {ok ? root : null_root, in.tri_base, in.mesh, 0u};
    std::memcpy(&rec[6], meta, sizeof(meta));
    rec[7] = make_float4(0.f, 0.f, 0.f, 0.f);
    if (!ok) {
        wb[0] = wb[1] = make_float4(in.xf[3], in.xf[7], in.xf[11], 0.f);
        return;
    }
    double lo[3] = {HUGE_VAL, HUGE_VAL, HUGE_VAL}, hi[3] = {-HUGE_VAL, -HUGE_VAL, -HUGE_VAL};
    for (int cn = 0; cn < 8; cn++) {
        const double p[3] = {box[(cn & 1) ? 3 : 0], box[(cn & 2) ? 4 : 1], box[(cn & 4) ? 5 : 2]};
        for (int r = 0; r < 3; r++) {
            const double w = m[r][0] * p[0] + m[r][1] * p[1] + m[r][2] * p[2] + t[r];
            lo[r] = std::min(lo[r], w);
            hi[r] = std::max(hi[r], w);
        }
    }
    double big = 0.0;
    for (int r = 0; r < 3; r++) big = std::max(big, std::max(std::fabs(lo[r]), std::fabs(hi[r])));
    const double pad = big * 0x1p-20 + 1e-30;
    wb[0] = make_float4((float)(lo[0] - pad), (float)(lo[1] - pad), (float)(lo[2] - pad), 0.f);
    wb[1] = make_float4((float)(hi[0] + pad), (float)(hi[1] + pad), (float)(hi[2] + pad), 0.f);
}

const float4 kFarBox = make_float4(1e38f, 1e38f, 1e38f, 0.f);   // the empty-slot box of the builder

// records and boxes of all instances (+ the null instance at ninst)
void tl_records(const rtx_ctx* c, const Built* b, std::vector<float4>& xrec, std::vector<float4>& wbox) {
    const uint32_t ninst = (uint32_t)c->things.size();
    xrec.assign((size_t)rtx::kInstF4 * (ninst + 1), make_float4(0.f, 0.f, 0.f, 0.f));
    wbox.assign(2 * (size_t)(ninst + 1), kFarBox);
    for (uint32_t i = 0; i < ninst; i++) {
        const uint32_t m = b->ib[i].mesh;
        float4* rec = &xrec[(size_t)rtx::kInstF4 * i];
        if (c->meshes[m].ni == 0) {
            // no triangles: not in the top level; a record for completeness
            float4 wb[2];
            tl_instance(b->ib[i], b->null_root, b->null_root, &b->blas_box[6 * (size_t)m], rec, wb);
            continue;
        }
        tl_instance(b->ib[i], b->blas_root[m], b->null_root, &b->blas_box[6 * (size_t)m], rec, &wbox[2 * (size_t)i]);
    }
    rtx::InstBuild id{};
    std::memcpy(id.xf, kIdentity, sizeof(id.xf));
    float4 wb[2];
    const float zero_box[6] = {0.f, 0.f, 0.f, 0.f, 0.f, 0.f};
    tl_instance(id, b->null_root, b->null_root, zero_box, &xrec[(size_t)rtx::kInstF4 * ninst], wb);
}

}  // namespace

// Builds a two-level scene into a new Built (not registered): one tree per
// mesh in use over its object-space triangles, one over the instances.
static rtx_status build_scene_tl(rtx_ctx* c, hipStream_t s, Built** out) {
    HIP_OK(hipSetDevice(c->device));
    const uint32_t ninst = (uint32_t)c->things.size();
    if (ninst >= rtx::kLeafFirstMask) return fail(RTX_ERROR_LIMIT, "rtx_scene_build: 2^27-1 or more instances");
    std::vector<rtx::InstBuild> ib;
    std::vector<float4> irec;
    uint64_t N = 0;
    instance_records(c, ib, irec, N);
    if (N >= 0xFFFFFFFFull) return fail(RTX_ERROR_LIMIT, "rtx_scene_build: 2^32-1 or more instanced triangles");
    const size_t nm = c->meshes.size();
    std::vector<char> used(nm, 0);
    for (uint32_t i = 0; i < ninst; i++)
        if (c->meshes[c->things[i].mesh].ni) used[c->things[i].mesh] = 1;

    hipEvent_t e0, e1;
    HIP_OK(hipEventCreate(&e0));
    HIP_OK(hipEventCreate(&e1));
    HIP_OK(hipEventRecord(e0, s));
    Built* b = new Built();
    b->device = c->device;
    b->two_level = true;
    b->ib = ib;
    auto bail = [&](rtx_status st) { b->release(); delete b; (void)hipEventDestroy(e0); (void)hipEventDestroy(e1); return st; };
    hipError_t e;
    std::string err;

    // bottom level: one tree per mesh in use
    BvhSet bl;
    bl.v.resize(nm);
    b->blas_root.assign(nm, rtx::kEmptyRoot);
    b->blas_box.assign(6 * (nm ? nm : 1), 0.f);
    {
        uint32_t nmax = 0;
        for (size_t m = 0; m < nm; m++)
            if (used[m]) nmax = std::max(nmax, c->meshes[m].ni);
        std::vector<rtx::MeshRef> mr(nm ? nm : 1);
        for (size_t m = 0; m < nm; m++)
            mr[m] = rtx::MeshRef{c->meshes[m].d_v, c->meshes[m].d_i, c->meshes[m].nv, c->meshes[m].ni};
        Tmp<rtx::InstBuild> d_one;
        Tmp<rtx::MeshRef> d_mr;
        Tmp<float4> d_tw, d_plo, d_phi;
        if ((e = hipMalloc((void**)&d_one.p, sizeof(rtx::InstBuild) * nm + sizeof(rtx::InstBuild))) != hipSuccess) return bail(hip_fail(e, "hipMalloc(insts)"));
        if ((e = hipMalloc((void**)&d_mr.p, sizeof(rtx::MeshRef) * mr.size())) != hipSuccess) return bail(hip_fail(e, "hipMalloc(meshes)"));
        if ((e = hipMalloc((void**)&d_tw.p, sizeof(float4) * 3 * (nmax ? nmax : 1))) != hipSuccess) return bail(hip_fail(e, "hipMalloc(triangles)"));
        if ((e = hipMalloc((void**)&d_plo.p, sizeof(float4) * (nmax ? nmax : 1))) != hipSuccess) return bail(hip_fail(e, "hipMalloc(bounds)"));
        if ((e = hipMalloc((void**)&d_phi.p, sizeof(float4) * (nmax ? nmax : 1))) != hipSuccess) return bail(hip_fail(e, "hipMalloc(bounds)"));
        if ((e = hipMemcpyAsync(d_mr.p, mr.data(), sizeof(rtx::MeshRef) * mr.size(), hipMemcpyHostToDevice, s)) != hipSuccess) return bail(hip_fail(e, "upload meshes"));
        // one identity instance per mesh (object space), uploaded at once
        std::vector<rtx::InstBuild> one(nm ? nm : 1);
        for (size_t m = 0; m < nm; m++) {
            std::memcpy(one[m].xf, kIdentity, sizeof(one[m].xf));
            one[m].mesh = (uint32_t)m;
            one[m].tri_base = 0;
        }
        if ((e = hipMemcpyAsync(d_one.p, one.data(), sizeof(rtx::InstBuild) * one.size(), hipMemcpyHostToDevice, s)) != hipSuccess) return bail(hip_fail(e, "upload insts"));
        for (size_t m = 0; m < nm; m++) {
            if (!used[m]) continue;
            const uint32_t ni = c->meshes[m].ni;
            if ((e = rtx::flatten_scene(d_one.p + m, 1, d_mr.p, ni, d_tw.p, d_plo.p, d_phi.p, s)) != hipSuccess) return bail(hip_fail(e, "flatten"));
            if ((e = rtx::build_bvh(d_tw.p, d_plo.p, d_phi.p, ni, s, &bl.v[m], &err)) != hipSuccess)
                return bail(fail(e == hipErrorOutOfMemory ? RTX_ERROR_OUT_OF_MEMORY : RTX_ERROR_HIP, "build_bvh: " + err));
            for (int k = 0; k < 3; k++) {
                b->blas_box[6 * m + k] = bl.v[m].root_lo[k];
                b->blas_box[6 * m + 3 + k] = bl.v[m].root_hi[k];
            }
        }
    }

    // scene arrays: top-level nodes first, then each mesh's tree; each
    // mesh's triangles with its degenerate one
    std::vector<uint64_t> tri_base(nm, 0), node_off(nm, 0);
    uint64_t tri_total = 0, blas_nodes = 0;
    uint32_t blas_depth = 0, blas_wide_depth = 0, max_leaf = 0, num_nodes = 0, num_leaves = 0;
    int first_used = -1;
    for (size_t m = 0; m < nm; m++) {
        if (!used[m]) continue;
        if (first_used < 0) first_used = (int)m;
        tri_base[m] = tri_total;
        tri_total += (uint64_t)c->meshes[m].ni + 1;
        node_off[m] = blas_nodes;
        blas_nodes += bl.v[m].num_wide;
        blas_depth = std::max(blas_depth, bl.v[m].max_depth);
        blas_wide_depth = std::max(blas_wide_depth, bl.v[m].wide_depth);
        max_leaf = std::max(max_leaf, bl.v[m].max_leaf);
        num_nodes += bl.v[m].num_nodes;
        num_leaves += bl.v[m].num_leaves;
    }
    if (tri_total > rtx::kLeafFirstMask || tri_total * 48u >= (1ull << 32))
        return bail(fail(RTX_ERROR_LIMIT, "rtx_scene_build: mesh triangles above 2^27-1 or 4 GiB"));
    b->null_root = first_used >= 0 ? rtx::leaf_ref((uint32_t)(tri_base[first_used] + c->meshes[first_used].ni), 1u)
                                   : rtx::kEmptyRoot;

    // top level over the instance boxes (instances of meshes with triangles)
    std::vector<float4> xrec, wbox;
    {
        // roots are needed for the records only, boxes now
        b->blas_root.assign(nm, b->null_root);
        tl_records(c, b, xrec, wbox);
    }
    std::vector<uint32_t> tl_inst;
    std::vector<float4> tplo, tphi;
    for (uint32_t i = 0; i < ninst; i++) {
        if (!c->meshes[c->things[i].mesh].ni) continue;
        tl_inst.push_back(i);
        tplo.push_back(wbox[2 * (size_t)i]);
        tphi.push_back(wbox[2 * (size_t)i + 1]);
    }
    const uint32_t M = (uint32_t)tl_inst.size();
    BvhSet tl;
    tl.v.resize(1);
    rtx::BuiltBVH& tb = tl.v[0];
    Tmp<uint32_t> d_inst_of;
    if (M) {
        Tmp<float4> d_tplo, d_tphi;
        if ((e = hipMalloc((void**)&d_tplo.p, sizeof(float4) * M)) != hipSuccess) return bail(hip_fail(e, "hipMalloc(instance boxes)"));
        if ((e = hipMalloc((void**)&d_tphi.p, sizeof(float4) * M)) != hipSuccess) return bail(hip_fail(e, "hipMalloc(instance boxes)"));
        if ((e = hipMalloc((void**)&d_inst_of.p, sizeof(uint32_t) * M)) != hipSuccess) return bail(hip_fail(e, "hipMalloc(instance ids)"));
        if ((e = hipMemcpyAsync(d_tplo.p, tplo.data(), sizeof(float4) * M, hipMemcpyHostToDevice, s)) != hipSuccess) return bail(hip_fail(e, "upload boxes"));
        if ((e = hipMemcpyAsync(d_tphi.p, tphi.data(), sizeof(float4) * M, hipMemcpyHostToDevice, s)) != hipSuccess) return bail(hip_fail(e, "upload boxes"));
        if ((e = hipMemcpyAsync(d_inst_of.p, tl_inst.data(), sizeof(uint32_t) * M, hipMemcpyHostToDevice, s)) != hipSuccess) return bail(hip_fail(e, "upload ids"));
        rtx::BvhOptions opt;
        opt.max_leaf = 1;
        opt.gather = false;
        if ((e = rtx::build_bvh(nullptr, d_tplo.p, d_tphi.p, M, s, &tb, &err, opt)) != hipSuccess)
            return bail(fail(e == hipErrorOutOfMemory ? RTX_ERROR_OUT_OF_MEMORY : RTX_ERROR_HIP, "build_bvh (top level): " + err));
    }
    const uint64_t total_nodes = (uint64_t)tb.num_wide + blas_nodes;
    if (total_nodes * 16u * rtx::kWideNodeF4 >= (1ull << 32))
        return bail(fail(RTX_ERROR_LIMIT, "rtx_scene_build: node array above 4 GiB"));

    if ((e = hipMalloc((void**)&b->nodes, sizeof(float4) * rtx::kWideNodeF4 * (total_nodes ? total_nodes : 1))) != hipSuccess) return bail(hip_fail(e, "hipMalloc(nodes)"));
    if ((e = hipMalloc((void**)&b->tris, sizeof(float4) * 3 * (tri_total ? tri_total : 1))) != hipSuccess) return bail(hip_fail(e, "hipMalloc(triangles)"));
    if (tb.num_wide) {
        if ((e = hipMemcpyAsync(b->nodes, tb.nodes, sizeof(float4) * rtx::kWideNodeF4 * tb.num_wide, hipMemcpyDeviceToDevice, s)) != hipSuccess) return bail(hip_fail(e, "copy top level"));
        if ((e = rtx::tlas_leaves(b->nodes, tb.num_wide, tb.order, M, d_inst_of.p, ninst, s)) != hipSuccess) return bail(hip_fail(e, "top-level leaves"));
    }
    for (size_t m = 0; m < nm; m++) {
        if (!used[m]) continue;
        const rtx::BuiltBVH& x = bl.v[m];
        const uint32_t nb = (uint32_t)(tb.num_wide + node_off[m]), tbm = (uint32_t)tri_base[m];
        if (x.num_wide) {
            if ((e = hipMemcpyAsync(b->nodes + (size_t)rtx::kWideNodeF4 * nb, x.nodes, sizeof(float4) * rtx::kWideNodeF4 * x.num_wide, hipMemcpyDeviceToDevice, s)) != hipSuccess) return bail(hip_fail(e, "copy bottom level"));
            if ((e = rtx::relocate_blas(b->nodes, x.num_wide, nb, tbm, s)) != hipSuccess) return bail(hip_fail(e, "relocate"));
        }
        if ((e = hipMemcpyAsync(b->tris + 3 * (size_t)tbm, x.tris, sizeof(float4) * 3 * ((size_t)c->meshes[m].ni + 1), hipMemcpyDeviceToDevice, s)) != hipSuccess) return bail(hip_fail(e, "copy triangles"));
        b->blas_root[m] = (x.root & rtx::kLeafFlag) ? x.root + tbm : x.root + nb;
    }
    // records with the roots, then the top level's boxes from them
    tl_records(c, b, xrec, wbox);
    if ((e = hipMalloc((void**)&b->xinst, sizeof(float4) * xrec.size())) != hipSuccess) return bail(hip_fail(e, "hipMalloc(instance records)"));
    if ((e = hipMalloc((void**)&b->ibox, sizeof(float4) * wbox.size())) != hipSuccess) return bail(hip_fail(e, "hipMalloc(instance boxes)"));
    if ((e = hipMalloc((void**)&b->insts, sizeof(float4) * irec.size())) != hipSuccess) return bail(hip_fail(e, "hipMalloc(inst records)"));
    if ((e = hipMemcpyAsync(b->xinst, xrec.data(), sizeof(float4) * xrec.size(), hipMemcpyHostToDevice, s)) != hipSuccess) return bail(hip_fail(e, "upload instance records"));
    if ((e = hipMemcpyAsync(b->ibox, wbox.data(), sizeof(float4) * wbox.size(), hipMemcpyHostToDevice, s)) != hipSuccess) return bail(hip_fail(e, "upload instance boxes"));
    if ((e = hipMemcpyAsync(b->insts, irec.data(), sizeof(float4) * irec.size(), hipMemcpyHostToDevice, s)) != hipSuccess) return bail(hip_fail(e, "upload inst records"));
    b->tlas_levels = tb.level_begin;
    if (tb.num_wide) {
        if ((e = rtx::refit_tlas(b->nodes, b->tlas_levels, b->ibox, rtx::kInstLeaf | ninst, s)) != hipSuccess) return bail(hip_fail(e, "refit"));
    }

    uint32_t root = rtx::kEmptyRoot;
    if (M == 1) root = rtx::kInstLeaf | tl_inst[0];
    else if (M > 1) root = tb.root;
    b->hdr.nodes = b->nodes;
    b->hdr.tris = b->tris;
    b->hdr.insts = b->insts;
    b->hdr.xinst = b->xinst;
    b->hdr.root = root;
    b->hdr.num_prims = (uint32_t)N;
    b->hdr.num_nodes = (uint32_t)total_nodes;
    b->hdr.num_insts = ninst;
    if ((e = hipMalloc((void**)&b->d_hdr, sizeof(rtx::DevScene))) != hipSuccess) return bail(hip_fail(e, "hipMalloc(handle)"));
    if ((e = hipMemcpyAsync(b->d_hdr, &b->hdr, sizeof(rtx::DevScene), hipMemcpyHostToDevice, s)) != hipSuccess) return bail(hip_fail(e, "upload handle"));
    if ((e = hipEventRecord(e1, s)) != hipSuccess) return bail(hip_fail(e, "hipEventRecord"));
    if ((e = hipStreamSynchronize(s)) != hipSuccess) return bail(hip_fail(e, "hipStreamSynchronize"));
    float ms = 0.f;
    (void)hipEventElapsedTime(&ms, e0, e1);
    (void)hipEventDestroy(e0);
    (void)hipEventDestroy(e1);
    if (std::max(blas_depth, tb.max_depth) > (uint32_t)rtx::kMaxTreeDepth) {
        b->release();
        delete b;
        return fail(RTX_ERROR_LIMIT, "rtx_scene_build: tree deeper than the traversal stack");
    }
    uint64_t unique = 0;
    for (size_t m = 0; m < nm; m++)
        if (used[m]) unique += c->meshes[m].ni;
    b->stats.num_prims = (uint32_t)N;
    b->stats.num_nodes = num_nodes + tb.num_nodes;
    b->stats.num_leaves = num_leaves + tb.num_leaves;
    b->stats.max_depth = tb.max_depth + blas_depth;
    b->stats.max_leaf = max_leaf;
    b->stats.sah_cost = (float)tb.sah;
    b->stats.build_ms = ms;
    b->stats.device_bytes = sizeof(float4) * ((uint64_t)rtx::kWideNodeF4 * total_nodes + 3ull * tri_total + irec.size() +
                                              xrec.size() + wbox.size());
    b->stats.wide_nodes = (uint32_t)total_nodes;
    b->stats.wide_depth = tb.wide_depth + blas_wide_depth;
    b->stats.two_level = 1;
    b->stats.unique_prims = (uint32_t)unique;
    b->stats.tlas_nodes = tb.num_wide;
    uint32_t nblas = 0;
    for (size_t m = 0; m < nm; m++) nblas += used[m] ? 1u : 0u;
    b->stats.num_blas = nblas;
    // top-level steps, the world ray and marker, bottom-level steps
    b->stack_need = 3u * tb.wide_depth + 7u + 3u * (blas_wide_depth > 0 ? blas_wide_depth : 1u);
    *out = b;
    return RTX_SUCCESS;
}

// Scene::update of a two-level scene: new instance records and boxes from
// the current transforms, and a refit of the top level (same topology, as
// the reference's OPTIX_BUILD_OPERATION_UPDATE, optx/scene.cxx:275-294)
static rtx_status refit_scene_tl(rtx_ctx* c, Built* b, hipStream_t s) {
    HIP_OK(hipSetDevice(c->device));
    const uint32_t ninst = (uint32_t)c->things.size();
    std::vector<float4> irec, xrec, wbox;
    uint64_t N = 0;
    instance_records(c, b->ib, irec, N);
    tl_records(c, b, xrec, wbox);
    hipEvent_t e0, e1;
    HIP_OK(hipEventCreate(&e0));
    HIP_OK(hipEventCreate(&e1));
    hipError_t e = hipEventRecord(e0, s);
    if (e == hipSuccess) e = hipMemcpyAsync(b->xinst, xrec.data(), sizeof(float4) * xrec.size(), hipMemcpyHostToDevice, s);
    if (e == hipSuccess) e = hipMemcpyAsync(b->ibox, wbox.data(), sizeof(float4) * wbox.size(), hipMemcpyHostToDevice, s);
    if (e == hipSuccess) e = hipMemcpyAsync(b->insts, irec.data(), sizeof(float4) * irec.size(), hipMemcpyHostToDevice, s);
    if (e == hipSuccess && b->tlas_levels.size() > 1) e = rtx::refit_tlas(b->nodes, b->tlas_levels, b->ibox, rtx::kInstLeaf | ninst, s);
    if (e == hipSuccess) e = hipEventRecord(e1, s);
    if (e == hipSuccess) e = hipStreamSynchronize(s);
    float ms = 0.f;
    if (e == hipSuccess) (void)hipEventElapsedTime(&ms, e0, e1);
    (void)hipEventDestroy(e0);
    (void)hipEventDestroy(e1);
    if (e != hipSuccess) return hip_fail(e, "rtx_scene_update (refit)");
    b->stats.build_ms = ms;
    return RTX_SUCCESS;
}

static rtx_status build_scene(rtx_ctx* c, hipStream_t s, Built** out) {
    return c->build_mode == RTX_BUILD_TWO_LEVEL ? build_scene_tl(c, s, out) : build_scene_flat(c, s, out);
}

rtx_status rtx_scene_set_build_mode(rtx_ctx* c, int mode) {
    if (!c) return fail(RTX_ERROR_INVALID_VALUE, "rtx_scene_set_build_mode: ctx is NULL");
    if (mode != RTX_BUILD_FLAT && mode != RTX_BUILD_TWO_LEVEL)
        return fail(RTX_ERROR_INVALID_VALUE, "rtx_scene_set_build_mode: unknown mode");
    c->build_mode = mode;
    return RTX_SUCCESS;
}

rtx_status rtx_scene_get_build_mode(rtx_ctx* c, int* mode) {
    if (!c || !mode) return fail(RTX_ERROR_INVALID_VALUE, "rtx_scene_get_build_mode: NULL argument");
    *mode = c->build_mode;
    return RTX_SUCCESS;
}

// Scene::build, optx/scene.cxx:227-273 (the previous build of this context is freed, as Scene::free does)
rtx_status rtx_scene_build(rtx_ctx* c, void* stream, uint64_t* is_handle) {
    if (!c || !is_handle) return fail(RTX_ERROR_INVALID_VALUE, "rtx_scene_build: NULL argument");
    Built* b = nullptr;
    rtx_status st = build_scene(c, (hipStream_t)stream, &b);
    if (st != RTX_SUCCESS) return st;
    drop_built(c);
    c->built = b;
    {
        std::lock_guard<std::mutex> lk(g_reg_mu);
        g_registry[(uint64_t)(uintptr_t)b->d_hdr] = b;
    }
    *is_handle = (uint64_t)(uintptr_t)b->d_hdr;
    return RTX_SUCCESS;
}

// Scene::update, optx/scene.cxx:275-294: brings the structure up to date with
// the current transforms (here a rebuild) and keeps the handle: the new scene
// header is written into the handle's device allocation.
rtx_status rtx_scene_update(rtx_ctx* c, uint64_t is_handle, void* stream) {
    if (!c) return fail(RTX_ERROR_INVALID_VALUE, "rtx_scene_update: ctx is NULL");
    if (!c->built || (uint64_t)(uintptr_t)c->built->d_hdr != is_handle)
        return fail(RTX_ERROR_INVALID_HANDLE, "rtx_scene_update: handle is not this context's scene");
    hipStream_t s = (hipStream_t)stream;
    // a two-level scene with the same instances refits its top level in place
    if (c->built->two_level && c->built->hdr.num_insts == (uint32_t)c->things.size())
        return refit_scene_tl(c, c->built, s);
    Built* b = nullptr;
    rtx_status st = c->built->two_level ? build_scene_tl(c, s, &b) : build_scene_flat(c, s, &b);
    if (st != RTX_SUCCESS) return st;
    Built* old = c->built;
    HIP_OK(hipMemcpyAsync(old->d_hdr, &b->hdr, sizeof(rtx::DevScene), hipMemcpyHostToDevice, s));
    HIP_OK(hipStreamSynchronize(s));
    (void)hipFree(b->d_hdr);
    b->d_hdr = old->d_hdr;
    old->d_hdr = nullptr;
    old->release();
    delete old;
    c->built = b;
    {
        std::lock_guard<std::mutex> lk(g_reg_mu);
        g_registry[is_handle] = b;
    }
    return RTX_SUCCESS;
}

rtx_status rtx_scene_stats_get(rtx_ctx* c, uint64_t is_handle, rtx_scene_stats* st) {
    if (!c || !st) return fail(RTX_ERROR_INVALID_VALUE, "rtx_scene_stats_get: NULL argument");
    Built* b = lookup(is_handle);
    if (!b) return fail(RTX_ERROR_INVALID_HANDLE, "rtx_scene_stats_get: unknown handle");
    *st = b->stats;
    return RTX_SUCCESS;
}

rtx_status rtx_scene_triangles(rtx_ctx* c, uint64_t is_handle, float* tris, uint32_t* tri_inst) {
    if (!c) return fail(RTX_ERROR_INVALID_VALUE, "rtx_scene_triangles: ctx is NULL");
    Built* b = lookup(is_handle);
    if (!b) return fail(RTX_ERROR_INVALID_HANDLE, "rtx_scene_triangles: unknown handle");
    const uint32_t N = b->hdr.num_prims;
    if (N == 0) return RTX_SUCCESS;
    HIP_OK(hipSetDevice(b->device));
    std::vector<float4> h(3 * (size_t)N);
    if (b->two_level) {
        // a two-level scene keeps no world triangles: flatten them for the copy
        const uint32_t ninst = b->hdr.num_insts;
        std::vector<rtx::MeshRef> mr(c->meshes.size() ? c->meshes.size() : 1);
        for (size_t m = 0; m < c->meshes.size(); m++)
            mr[m] = rtx::MeshRef{c->meshes[m].d_v, c->meshes[m].d_i, c->meshes[m].nv, c->meshes[m].ni};
        Tmp<rtx::InstBuild> d_ib;
        Tmp<rtx::MeshRef> d_mr;
        Tmp<float4> d_tw, d_plo, d_phi;
        HIP_OK(hipMalloc((void**)&d_ib.p, sizeof(rtx::InstBuild) * b->ib.size()));
        HIP_OK(hipMalloc((void**)&d_mr.p, sizeof(rtx::MeshRef) * mr.size()));
        HIP_OK(hipMalloc((void**)&d_tw.p, sizeof(float4) * h.size()));
        HIP_OK(hipMalloc((void**)&d_plo.p, sizeof(float4) * N));
        HIP_OK(hipMalloc((void**)&d_phi.p, sizeof(float4) * N));
        HIP_OK(hipMemcpy(d_ib.p, b->ib.data(), sizeof(rtx::InstBuild) * b->ib.size(), hipMemcpyHostToDevice));
        HIP_OK(hipMemcpy(d_mr.p, mr.data(), sizeof(rtx::MeshRef) * mr.size(), hipMemcpyHostToDevice));
        HIP_OK(rtx::flatten_scene(d_ib.p, ninst, d_mr.p, N, d_tw.p, d_plo.p, d_phi.p, nullptr));
        HIP_OK(hipMemcpy(h.data(), d_tw.p, sizeof(float4) * h.size(), hipMemcpyDeviceToHost));
    } else {
        HIP_OK(hipMemcpy(h.data(), b->tri_world, sizeof(float4) * h.size(), hipMemcpyDeviceToHost));
    }
    for (size_t p = 0; p < N; p++) {
        for (int k = 0; k < 3; k++) {
            const float4 v = h[3 * p + k];
            if (tris) { tris[9 * p + 3 * k] = v.x; tris[9 * p + 3 * k + 1] = v.y; tris[9 * p + 3 * k + 2] = v.z; }
        }
        if (tri_inst) std::memcpy(&tri_inst[p], &h[3 * p + 1].w, 4);
    }
    return RTX_SUCCESS;
}

// ---- launch: Launcher::ignite, optx/launcher.cxx:49-77 ---------------------
rtx_status rtx_launch(rtx_ctx* c, const rtx_lp_general* lp, int once, void* stream) {
    if (!c || !lp) return fail(RTX_ERROR_INVALID_VALUE, "rtx_launch: NULL argument");
    Built* b = lookup(lp->is_handle);
    if (!b) return fail(RTX_ERROR_INVALID_HANDLE, "rtx_launch: unknown is_handle");
    if (b->device != c->device) return fail(RTX_ERROR_INVALID_HANDLE, "rtx_launch: handle is on another device");
    if (lp->image_w == 0 || lp->image_h == 0) return fail(RTX_ERROR_INVALID_VALUE, "rtx_launch: empty image");
    if ((uint64_t)lp->image_w * lp->image_h > 0xFFFFFFFFull / 2)
        return fail(RTX_ERROR_LIMIT, "rtx_launch: image too large");
    HIP_OK(hipSetDevice(c->device));
    if (rtx_status st = ensure_gstack(c->render_ls, b->stack_need)) return st;
    const float3 amb = make_float3(c->ambient.x, c->ambient.y, c->ambient.z);
    hipError_t e;
    if (once || lp->picker) {
        if (!lp->pick_id) return fail(RTX_ERROR_INVALID_VALUE, "rtx_launch: picker without pick_id");
        if (lp->pick_x >= lp->image_w || lp->pick_y >= lp->image_h)
            return fail(RTX_ERROR_INVALID_VALUE, "rtx_launch: pick position outside the image");
        e = rtx::launch_pick(b->hdr, *lp, c->render_ls, (hipStream_t)stream);
    } else {
        if (lp->spp == 0) return fail(RTX_ERROR_INVALID_VALUE, "rtx_launch: spp == 0");
        if (lp->depth == 0) return fail(RTX_ERROR_INVALID_VALUE, "rtx_launch: depth == 0");
        if (!lp->rawRGB) return fail(RTX_ERROR_INVALID_VALUE, "rtx_launch: rawRGB is NULL");
        if (rtx::frame_ordered(*lp, c->render_ls))
            if (rtx_status st = ensure_order(c->render_ls, rtx::frame_tiles(*lp, c->render_ls.tile_w))) return st;
        e = rtx::launch_render(b->hdr, *lp, amb, c->render_ls, (hipStream_t)stream);
    }
    if (e != hipSuccess) return hip_fail(e, "rtx_launch");
    return RTX_SUCCESS;
}

rtx_status rtx_launch_noalloc(rtx_ctx* c, const rtx_lp_general* lp, int once, void* stream) {
    if (!c || !lp) return fail(RTX_ERROR_INVALID_VALUE, "rtx_launch_noalloc: NULL argument");
    Built* b = lookup(lp->is_handle);
    if (!b) return fail(RTX_ERROR_INVALID_HANDLE, "rtx_launch_noalloc: unknown is_handle");
    if (b->device != c->device) return fail(RTX_ERROR_INVALID_HANDLE, "rtx_launch_noalloc: handle is on another device");
    if (lp->image_w == 0 || lp->image_h == 0) return fail(RTX_ERROR_INVALID_VALUE, "rtx_launch_noalloc: empty image");
    if ((uint64_t)lp->image_w * lp->image_h > 0xFFFFFFFFull / 2)
        return fail(RTX_ERROR_LIMIT, "rtx_launch_noalloc: image too large");
    uint8_t* base = nullptr;
    if (c->device < kNoAllocMaxDevices) {
        std::lock_guard<std::mutex> lk(g_scratch_mu);
        if (g_scratch_done[c->device] && !g_scratch_ptr[c->device])
            return hip_fail(g_scratch_err[c->device], "rtx_launch_noalloc: hipGetSymbolAddress(scratch)");
        base = g_scratch_ptr[c->device];
    }
    if (!base) return fail(RTX_ERROR_LIMIT, "rtx_launch_noalloc: device " + std::to_string(c->device) +
                                                " has no static scratch (devices 0.." +
                                                std::to_string(kNoAllocMaxDevices - 1) + " have)");
    // the launch's scratch inside the static array; the stack overflow area
    // covers the entries beyond the LDS part (the picker's LDS part is never smaller)
    rtx::LaunchScratch ls = c->render_ls;
    const uint32_t over = b->stack_need > ls.stack_lds ? b->stack_need - ls.stack_lds : 1u;
    const size_t stack_bytes = sizeof(uint32_t) * (size_t)over * ls.gstride;
    if (stack_bytes > kNoAllocBytes - kNoAllocStackOffset)
        return fail(RTX_ERROR_LIMIT, "rtx_launch_noalloc: the scene's traversal stack needs " +
                                         std::to_string(stack_bytes) + " bytes of overflow for " +
                                         std::to_string(ls.gstride) + " lanes, more than the " +
                                         std::to_string(kNoAllocBytes - kNoAllocStackOffset) + " of the static scratch");
    ls.counter = (uint32_t*)base;
    ls.tile_cost = (uint32_t*)(base + 256);
    ls.tile_order = ls.tile_cost + kNoAllocTiles;
    ls.tile_cap = kNoAllocTiles;
    ls.gstack = (uint32_t*)(base + kNoAllocStackOffset);
    ls.gstack_entries = over;
    HIP_OK(hipSetDevice(c->device));
    const float3 amb = make_float3(c->ambient.x, c->ambient.y, c->ambient.z);
    hipError_t e;
    if (once || lp->picker) {
        if (!lp->pick_id) return fail(RTX_ERROR_INVALID_VALUE, "rtx_launch_noalloc: picker without pick_id");
        if (lp->pick_x >= lp->image_w || lp->pick_y >= lp->image_h)
            return fail(RTX_ERROR_INVALID_VALUE, "rtx_launch_noalloc: pick position outside the image");
        e = rtx::launch_pick(b->hdr, *lp, ls, (hipStream_t)stream);
    } else {
        if (lp->spp == 0) return fail(RTX_ERROR_INVALID_VALUE, "rtx_launch_noalloc: spp == 0");
        if (lp->depth == 0) return fail(RTX_ERROR_INVALID_VALUE, "rtx_launch_noalloc: depth == 0");
        if (!lp->rawRGB) return fail(RTX_ERROR_INVALID_VALUE, "rtx_launch_noalloc: rawRGB is NULL");
        if (rtx::frame_ordered(*lp, ls) && rtx::frame_tiles(*lp, ls.tile_w) > kNoAllocTiles)
            return fail(RTX_ERROR_LIMIT, "rtx_launch_noalloc: a frame rendered in probe order has " +
                                             std::to_string(rtx::frame_tiles(*lp, ls.tile_w)) +
                                             " tiles, more than the " + std::to_string(kNoAllocTiles) +
                                             " of the static scratch");
        e = rtx::launch_render(b->hdr, *lp, amb, ls, (hipStream_t)stream);
    }
    if (e != hipSuccess) return hip_fail(e, "rtx_launch_noalloc");
    return RTX_SUCCESS;
}

// ---- batched render -----------------------------------------------------
}  // extern "C"

struct rtx_batch {
    rtx_ctx* ctx = nullptr;
    rtx::Job* d_jobs = nullptr;
    std::vector<rtx::Job> jobs;       // host copy of the job table
    std::vector<uint64_t> handles;    // scene of each job
    rtx::Job first{};
    uint32_t njobs = 0, nwork = 0, tile_w = 8;
    bool use_graph = false;
    hipGraph_t graph = nullptr;
    hipGraphExec_t exec = nullptr;
    rtx::LaunchScratch captured{};    // the scratch and tuning the graph's kernel arguments hold
};

static bool same_scratch(const rtx::LaunchScratch& a, const rtx::LaunchScratch& b) {
    return a.counter == b.counter && a.gstack == b.gstack && a.gstride == b.gstride && a.grid == b.grid &&
           a.shade_batch == b.shade_batch && a.inner_frac == b.inner_frac && a.stack_lds == b.stack_lds &&
           a.iterative == b.iterative;
}

extern "C" {

rtx_status rtx_batch_create(rtx_ctx* c, const rtx_lp_general* lps, uint32_t n, int use_graph, rtx_batch** out) {
    if (!c || !out || (n && !lps)) return fail(RTX_ERROR_INVALID_VALUE, "rtx_batch_create: NULL argument");
    if (n == 0) return fail(RTX_ERROR_INVALID_VALUE, "rtx_batch_create: empty batch");
    std::vector<uint32_t> order(n);
    std::vector<Built*> built(n);
    for (uint32_t i = 0; i < n; i++) {
        const rtx_lp_general& lp = lps[i];
        built[i] = lookup(lp.is_handle);
        if (!built[i]) return fail(RTX_ERROR_INVALID_HANDLE, "rtx_batch_create: unknown is_handle in frame " + std::to_string(i));
        if (built[i]->device != c->device) return fail(RTX_ERROR_INVALID_HANDLE, "rtx_batch_create: handle on another device");
        if (built[i]->two_level != built[0]->two_level)
            return fail(RTX_ERROR_INVALID_VALUE, "rtx_batch_create: flat and two-level scenes in one batch");
        if (lp.image_w == 0 || lp.image_h == 0 || lp.spp == 0 || lp.depth == 0 || !lp.rawRGB)
            return fail(RTX_ERROR_INVALID_VALUE, "rtx_batch_create: bad launch parameters in frame " + std::to_string(i));
        if ((uint64_t)lp.image_w * lp.image_h > 0xFFFFFFFFull / 2)
            return fail(RTX_ERROR_LIMIT, "rtx_batch_create: image too large in frame " + std::to_string(i));
        order[i] = i;
    }
    // longest work items first: a work item is one pixel whose cost grows
    // with spp x depth, so frames are ordered by that, then by size
    std::stable_sort(order.begin(), order.end(), [&](uint32_t a, uint32_t b) {
        const uint64_t ca = (uint64_t)lps[a].spp * lps[a].depth, cb = (uint64_t)lps[b].spp * lps[b].depth;
        if (ca != cb) return ca > cb;
        return (uint64_t)lps[a].image_w * lps[a].image_h > (uint64_t)lps[b].image_w * lps[b].image_h;
    });
    const uint32_t tw = c->render_ls.tile_w ? c->render_ls.tile_w : 8u;
    const float3 amb = make_float3(c->ambient.x, c->ambient.y, c->ambient.z);
    std::vector<rtx::Job> jobs(n);
    uint64_t work = 0;
    for (uint32_t k = 0; k < n; k++) {
        const uint32_t i = order[k];
        jobs[k] = rtx::make_job(built[i]->hdr, lps[i], amb, tw, (uint32_t)work);
        work = jobs[k].work_end;
        if (work >= 0xFFFFFFFFull) return fail(RTX_ERROR_LIMIT, "rtx_batch_create: more than 2^32 pixel slots");
    }
    HIP_OK(hipSetDevice(c->device));
    uint32_t need = 1;
    for (uint32_t i = 0; i < n; i++) need = std::max(need, built[i]->stack_need);
    if (rtx_status st = ensure_gstack(c->render_ls, need)) return st;
    rtx_batch* b = new rtx_batch();
    b->ctx = c;
    b->njobs = n;
    b->nwork = (uint32_t)work;
    b->tile_w = tw;
    b->first = jobs[0];
    b->jobs = jobs;
    b->handles.resize(n);
    for (uint32_t k = 0; k < n; k++) b->handles[k] = lps[order[k]].is_handle;
    b->use_graph = use_graph != 0;
    hipError_t e = hipMalloc((void**)&b->d_jobs, sizeof(rtx::Job) * n);
    if (e == hipSuccess) e = hipMemcpy(b->d_jobs, jobs.data(), sizeof(rtx::Job) * n, hipMemcpyHostToDevice);
    if (e != hipSuccess) {
        if (b->d_jobs) (void)hipFree(b->d_jobs);
        delete b;
        return hip_fail(e, "rtx_batch_create: job table");
    }
    *out = b;
    return RTX_SUCCESS;
}

rtx_status rtx_batch_launch(rtx_batch* b, void* stream) {
    if (!b) return fail(RTX_ERROR_INVALID_VALUE, "rtx_batch_launch: batch is NULL");
    HIP_OK(hipSetDevice(b->ctx->device));
    hipStream_t s = (hipStream_t)stream;
    // a scene rebuilt by rtx_scene_update since the job table was written
    // has new arrays: write the table again
    bool stale = false;
    uint32_t need = 1;
    for (uint32_t k = 0; k < b->njobs; k++) {
        Built* sc = lookup(b->handles[k]);
        if (!sc) return fail(RTX_ERROR_INVALID_HANDLE, "rtx_batch_launch: a frame's scene was destroyed");
        if (sc->two_level != (b->first.sc.xinst != nullptr))
            return fail(RTX_ERROR_INVALID_VALUE, "rtx_batch_launch: a frame's scene changed its build mode");
        need = std::max(need, sc->stack_need);
        if (std::memcmp(&b->jobs[k].sc, &sc->hdr, sizeof(rtx::DevScene)) != 0) {
            b->jobs[k].sc = sc->hdr;
            stale = true;
        }
    }
    if (stale) {
        HIP_OK(hipStreamSynchronize(s));
        HIP_OK(hipMemcpy(b->d_jobs, b->jobs.data(), sizeof(rtx::Job) * b->njobs, hipMemcpyHostToDevice));
        b->first = b->jobs[0];
        if (rtx_status st = ensure_gstack(b->ctx->render_ls, need)) return st;
        if (b->exec) {
            // the captured kernel arguments hold the first job: capture again
            (void)hipGraphExecDestroy(b->exec);
            (void)hipGraphDestroy(b->graph);
            b->exec = nullptr;
            b->graph = nullptr;
        }
    }
    if (!b->use_graph || !s) {
        hipError_t e = rtx::launch_render_batch(b->d_jobs, b->first, b->njobs, b->nwork, b->ctx->render_ls, b->tile_w, s);
        if (e != hipSuccess) return hip_fail(e, "rtx_batch_launch");
        return RTX_SUCCESS;
    }
    if (b->exec && !same_scratch(b->captured, b->ctx->render_ls)) {
        // the context's scratch moved (a deeper scene grew the stack area) or its
        // tuning changed since the capture: capture again
        HIP_OK(hipStreamSynchronize(s));
        (void)hipGraphExecDestroy(b->exec);
        (void)hipGraphDestroy(b->graph);
        b->exec = nullptr;
        b->graph = nullptr;
    }
    if (!b->exec) {
        b->captured = b->ctx->render_ls;
        HIP_OK(hipStreamBeginCapture(s, hipStreamCaptureModeThreadLocal));
        hipError_t e = rtx::launch_render_batch(b->d_jobs, b->first, b->njobs, b->nwork, b->ctx->render_ls, b->tile_w, s);
        hipGraph_t g = nullptr;
        hipError_t e2 = hipStreamEndCapture(s, &g);
        if (e != hipSuccess) { if (g) (void)hipGraphDestroy(g); return hip_fail(e, "rtx_batch_launch: capture"); }
        if (e2 != hipSuccess) return hip_fail(e2, "hipStreamEndCapture");
        b->graph = g;
        HIP_OK(hipGraphInstantiate(&b->exec, b->graph, nullptr, nullptr, 0));
    }
    HIP_OK(hipGraphLaunch(b->exec, s));
    return RTX_SUCCESS;
}

rtx_status rtx_batch_destroy(rtx_batch* b) {
    if (!b) return RTX_SUCCESS;
    (void)hipSetDevice(b->ctx->device);
    if (b->exec) (void)hipGraphExecDestroy(b->exec);
    if (b->graph) (void)hipGraphDestroy(b->graph);
    if (b->d_jobs) (void)hipFree(b->d_jobs);
    delete b;
    return RTX_SUCCESS;
}

// ---- differentiable rendering ----------------------------------------------
static Built* material_target(rtx_ctx* c, uint64_t h, const char* what) {
    Built* b = lookup(h);
    if (!b) {
        fail(RTX_ERROR_INVALID_HANDLE, std::string(what) + ": unknown handle");
        return nullptr;
    }
    if (b->device != c->device) {
        fail(RTX_ERROR_INVALID_HANDLE, std::string(what) + ": handle is on another device");
        return nullptr;
    }
    return b;
}

rtx_status rtx_scene_set_materials(rtx_ctx* c, uint64_t is_handle, const float* params, void* stream) {
    if (!c || !params) return fail(RTX_ERROR_INVALID_VALUE, "rtx_scene_set_materials: NULL argument");
    Built* b = material_target(c, is_handle, "rtx_scene_set_materials");
    if (!b) return RTX_ERROR_INVALID_HANDLE;
    HIP_OK(hipSetDevice(c->device));
    hipError_t e = rtx::launch_set_materials(b->hdr, params, (hipStream_t)stream);
    if (e != hipSuccess) return hip_fail(e, "rtx_scene_set_materials");
    return RTX_SUCCESS;
}

rtx_status rtx_scene_get_materials(rtx_ctx* c, uint64_t is_handle, float* params, void* stream) {
    if (!c || !params) return fail(RTX_ERROR_INVALID_VALUE, "rtx_scene_get_materials: NULL argument");
    Built* b = material_target(c, is_handle, "rtx_scene_get_materials");
    if (!b) return RTX_ERROR_INVALID_HANDLE;
    HIP_OK(hipSetDevice(c->device));
    hipError_t e = rtx::launch_get_materials(b->hdr, params, (hipStream_t)stream);
    if (e != hipSuccess) return hip_fail(e, "rtx_scene_get_materials");
    return RTX_SUCCESS;
}

rtx_status rtx_launch_adjoint(rtx_ctx* c, const rtx_lp_general* lp, const rtx_float3* dL_drgb, float* d_params,
                              void* stream) {
    if (!c || !lp || !dL_drgb || !d_params) return fail(RTX_ERROR_INVALID_VALUE, "rtx_launch_adjoint: NULL argument");
    Built* b = material_target(c, lp->is_handle, "rtx_launch_adjoint");
    if (!b) return RTX_ERROR_INVALID_HANDLE;
    if (lp->image_w == 0 || lp->image_h == 0 || lp->spp == 0 || lp->depth == 0 || !lp->rawRGB)
        return fail(RTX_ERROR_INVALID_VALUE, "rtx_launch_adjoint: bad launch parameters");
    if ((uint64_t)lp->image_w * lp->image_h > 0xFFFFFFFFull / 2)
        return fail(RTX_ERROR_LIMIT, "rtx_launch_adjoint: image too large");
    if (lp->depth > 64) return fail(RTX_ERROR_LIMIT, "rtx_launch_adjoint: depth > 64");
    if (b->hdr.num_insts >= (1u << 24)) return fail(RTX_ERROR_LIMIT, "rtx_launch_adjoint: 2^24 or more instances");
    HIP_OK(hipSetDevice(c->device));
    if (rtx_status st = ensure_gstack(c->render_ls, b->stack_need)) return st;
    if (c->tape_depth < lp->depth) {
        // path records of 14 words for every lane of the grid: a path records
        // each hit it shades, at most depth - 1 under the recursive programs
        // (the depth-th ray's hit is black unshaded) and depth under the
        // iterative ones (that hit is shaded too); sized for depth
        if (c->tape) {
            HIP_OK(hipDeviceSynchronize());
            HIP_OK(hipFree(c->tape));
            c->tape = nullptr;
            c->tape_depth = 0;
        }
        const size_t words = (size_t)lp->depth * rtx::kAdjointTapeWords * c->render_ls.gstride;
        HIP_OK(hipMalloc((void**)&c->tape, words * sizeof(uint32_t)));
        c->tape_depth = lp->depth;
    }
    if (rtx::frame_ordered(*lp, c->render_ls))
        if (rtx_status st = ensure_order(c->render_ls, rtx::frame_tiles(*lp, c->render_ls.tile_w))) return st;
    const float3 amb = make_float3(c->ambient.x, c->ambient.y, c->ambient.z);
    hipError_t e = rtx::launch_adjoint(b->hdr, *lp, amb, c->render_ls, dL_drgb, d_params, c->tape,
                                       c->render_ls.gstride, (hipStream_t)stream);
    if (e != hipSuccess) return hip_fail(e, "rtx_launch_adjoint");
    return RTX_SUCCESS;
}

rtx_status rtx_launch_counted(rtx_ctx* c, const rtx_lp_general* lp, uint64_t* node_visits, uint64_t* tri_tests,
                              uint64_t* wave_cycles, void* stream) {
    if (!c || !lp || !node_visits || !tri_tests) return fail(RTX_ERROR_INVALID_VALUE, "rtx_launch_counted: NULL argument");
    Built* b = lookup(lp->is_handle);
    if (!b) return fail(RTX_ERROR_INVALID_HANDLE, "rtx_launch_counted: unknown is_handle");
    if (lp->image_w == 0 || lp->image_h == 0 || lp->spp == 0 || lp->depth == 0 || !lp->rawRGB)
        return fail(RTX_ERROR_INVALID_VALUE, "rtx_launch_counted: bad launch parameters");
    HIP_OK(hipSetDevice(c->device));
    if (rtx_status st = ensure_gstack(c->render_ls, b->stack_need)) return st;
    rtx::TraceCounters tc;
    tc.nodes = node_visits;
    tc.tris = tri_tests;
    tc.wave_cycles = wave_cycles;
    const float3 amb = make_float3(c->ambient.x, c->ambient.y, c->ambient.z);
    hipError_t e = rtx::launch_render_counted(b->hdr, *lp, amb, c->render_ls, tc, (hipStream_t)stream);
    if (e != hipSuccess) return hip_fail(e, "rtx_launch_counted");
    return RTX_SUCCESS;
}

rtx_status rtx_trace(rtx_ctx* c, uint64_t is_handle, const rtx_ray* rays, uint32_t n, rtx_hit* hits, void* stream) {
    if (!c) return fail(RTX_ERROR_INVALID_VALUE, "rtx_trace: ctx is NULL");
    Built* b = lookup(is_handle);
    if (!b) return fail(RTX_ERROR_INVALID_HANDLE, "rtx_trace: unknown handle");
    if (n && (!rays || !hits)) return fail(RTX_ERROR_INVALID_VALUE, "rtx_trace: NULL buffer");
    HIP_OK(hipSetDevice(c->device));
    if (rtx_status st = ensure_gstack(c->trace_ls, b->stack_need)) return st;
    hipError_t e = rtx::launch_trace(b->hdr, rays, n, hits, c->trace_ls, (hipStream_t)stream);
    if (e != hipSuccess) return hip_fail(e, "rtx_trace");
    return RTX_SUCCESS;
}

rtx_status rtx_rng_draws(const uint32_t* seeds, uint32_t n, uint32_t k, float* out, void* stream) {
    if (n && k && (!seeds || !out)) return fail(RTX_ERROR_INVALID_VALUE, "rtx_rng_draws: NULL buffer");
    hipError_t e = rtx::launch_rng(seeds, n, k, out, (hipStream_t)stream);
    if (e != hipSuccess) return hip_fail(e, "rtx_rng_draws");
    return RTX_SUCCESS;
}

rtx_status rtx_pp_srgb(const rtx_float3* src, rtx_uchar4* dst, int w, int h, void* stream) {
    if (w < 0 || h < 0) return fail(RTX_ERROR_INVALID_VALUE, "rtx_pp_srgb: negative size");
    if ((w && h) && (!src || !dst)) return fail(RTX_ERROR_INVALID_VALUE, "rtx_pp_srgb: NULL buffer");
    hipError_t e = rtx::launch_pp(src, dst, w, h, true, (hipStream_t)stream);
    if (e != hipSuccess) return hip_fail(e, "rtx_pp_srgb");
    return RTX_SUCCESS;
}

rtx_status rtx_pp_none(const rtx_float3* src, rtx_uchar4* dst, int w, int h, void* stream) {
    if (w < 0 || h < 0) return fail(RTX_ERROR_INVALID_VALUE, "rtx_pp_none: negative size");
    if ((w && h) && (!src || !dst)) return fail(RTX_ERROR_INVALID_VALUE, "rtx_pp_none: NULL buffer");
    hipError_t e = rtx::launch_pp(src, dst, w, h, false, (hipStream_t)stream);
    if (e != hipSuccess) return hip_fail(e, "rtx_pp_none");
    return RTX_SUCCESS;
}

// ---- plumbing ---------------------------------------------------------------
rtx_status rtx_malloc(void** p, size_t bytes) {
    if (!p) return fail(RTX_ERROR_INVALID_VALUE, "rtx_malloc: NULL argument");
    HIP_OK(hipMalloc(p, bytes ? bytes : 1));
    return RTX_SUCCESS;
}
rtx_status rtx_free(void* p) {
    if (p) HIP_OK(hipFree(p));
    return RTX_SUCCESS;
}
rtx_status rtx_memcpy_htod(void* dst, const void* src, size_t bytes) {
    if (bytes) HIP_OK(hipMemcpy(dst, src, bytes, hipMemcpyHostToDevice));
    return RTX_SUCCESS;
}
rtx_status rtx_memcpy_dtoh(void* dst, const void* src, size_t bytes) {
    if (bytes) HIP_OK(hipMemcpy(dst, src, bytes, hipMemcpyDeviceToHost));
    return RTX_SUCCESS;
}
rtx_status rtx_memset(void* p, int value, size_t bytes) {
    if (bytes) HIP_OK(hipMemset(p, value, bytes));
    return RTX_SUCCESS;
}
rtx_status rtx_synchronize(void) {
    HIP_OK(hipDeviceSynchronize());
    return RTX_SUCCESS;
}
rtx_status rtx_device_count(int* n) {
    if (!n) return fail(RTX_ERROR_INVALID_VALUE, "rtx_device_count: NULL argument");
    if (hipGetDeviceCount(n) != hipSuccess) *n = 0;
    return RTX_SUCCESS;
}

}  // extern "C"
