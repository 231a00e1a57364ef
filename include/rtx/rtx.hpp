// rtx.hpp -- C++ host classes with the reference's hot-path API, over the
// C-ABI of librtx.so (include/rtx.h).  Header-only; link with -lrtx.
//
//   Scene     optx/scene.h:19-63      load/add/set/get/build/update/operator[]
//   Launcher  optx/launcher.h:17-31   resize/ignite(stream, once)
//   Camera    optx/camera.h:28-94     set + setters that re-run set
//   Sphere    optx/sphere.h:20-38     Sphere(radius, ndiv).mesh()
//   Object    optx/object.h:20-33     list of submeshes
//   Thing/Optics/LpGeneral            optx/thing.h:17-52, optx/rtwo.h:26-52
//   pp_sRGB/pp_none                   optx/postproc.cu:49-61
//
// Failures throw std::runtime_error("rtx error: <message>"), as CUDA_CHECK /
// OPTX_CHECK do in the reference (optx/util_cpu.h:18-53).
#pragma once

#include <cstring>
#include <stdexcept>
#include <string>
#include <tuple>
#include <vector>

#include "rtx.h"

namespace rtx {

inline void check(rtx_status st, const char* what) {
    if (st != RTX_SUCCESS) throw std::runtime_error(std::string("rtx error: ") + what + ": " + rtx_last_error());
}

typedef rtx_float3 float3;
typedef rtx_uint3 uint3;
typedef rtx_uchar4 uchar4;
typedef rtx_optics Optics;
typedef rtx_thing Thing;
typedef rtx_lp_general LpGeneral;
typedef uint64_t TraversableHandle;

// Mesh = tuple<float3*, uint, uint3*, uint> (optx/sphere.h:18)
typedef std::tuple<float3*, unsigned int, uint3*, unsigned int> Mesh;

class Sphere {
   public:
    explicit Sphere(float radius = 1.f, unsigned int ndiv = 6) {
        uint32_t nv = 0, ni = 0;
        check(rtx_sphere_mesh(radius, ndiv, nullptr, 0, &nv, nullptr, 0, &ni), "rtx_sphere_mesh");
        vces_.resize(nv);
        ices_.resize(ni);
        check(rtx_sphere_mesh(radius, ndiv, vces_.data(), nv, &nv, ices_.data(), ni, &ni), "rtx_sphere_mesh");
        vces_.resize(nv);
    }
    Mesh mesh() { return Mesh(vces_.data(), (unsigned)vces_.size(), ices_.data(), (unsigned)ices_.size()); }

   private:
    std::vector<float3> vces_;
    std::vector<uint3> ices_;
};

class Object {
   public:
    Object() {}
    // Object(wavefront), optx/object.cxx:23-93: one submesh per OBJ shape
    explicit Object(const std::string& wavefront) {
        rtx_object* o = nullptr;
        check(rtx_object_load(wavefront.c_str(), &o), "rtx_object_load");
        unsigned int n = 0;
        rtx_object_size(o, &n);
        for (unsigned int m = 0; m < n; m++) {
            const float3* v;
            const uint3* i;
            unsigned int nv, ni;
            rtx_object_mesh(o, m, &v, &nv, &i, &ni);
            vces_.emplace_back(v, v + nv);
            ices_.emplace_back(i, i + ni);
        }
        rtx_object_free(o);
    }
    void push(const std::vector<float3>& v, const std::vector<uint3>& i) {
        vces_.push_back(v);
        ices_.push_back(i);
    }
    Mesh operator[](unsigned int m) {
        return Mesh(vces_[m].data(), (unsigned)vces_[m].size(), ices_[m].data(), (unsigned)ices_[m].size());
    }
    size_t size() const { return vces_.size(); }

   private:
    std::vector<std::vector<float3>> vces_;
    std::vector<std::vector<uint3>> ices_;
};

class Camera {
   public:
    void set(const float3& eye, const float3& pat, const float3& vup, float fov, float aspratio, float aperture,
             float fostance) {
        check(rtx_camera_set(&c_, eye, pat, vup, fov, aspratio, aperture, fostance), "rtx_camera_set");
    }
    float3 eye() const { return c_.eye; }
    void eye(const float3& v) { set(v, c_.pat, c_.vup, c_.fov, c_.aspratio, c_.aperture, c_.fostance); }
    float3 pat() const { return c_.pat; }
    void pat(const float3& v) { set(c_.eye, v, c_.vup, c_.fov, c_.aspratio, c_.aperture, c_.fostance); }
    float3 vup() const { return c_.vup; }
    void vup(const float3& v) { set(c_.eye, c_.pat, v, c_.fov, c_.aspratio, c_.aperture, c_.fostance); }
    float fov() const { return c_.fov; }
    void fov(float v) { set(c_.eye, c_.pat, c_.vup, v, c_.aspratio, c_.aperture, c_.fostance); }
    void aspratio(float v) { set(c_.eye, c_.pat, c_.vup, c_.fov, v, c_.aperture, c_.fostance); }
    float aperture() const { return c_.aperture; }
    void aperture(float v) { set(c_.eye, c_.pat, c_.vup, c_.fov, c_.aspratio, v, c_.fostance); }
    float fostance() const { return c_.fostance; }
    void fostance(float v) { set(c_.eye, c_.pat, c_.vup, c_.fov, c_.aspratio, c_.aperture, v); }
    const rtx_camera& pod() const { return c_; }

   private:
    rtx_camera c_{};
};

class Scene {
   public:
    explicit Scene(int device = 0) { check(rtx_ctx_create(device, &ctx_), "rtx_ctx_create"); }
    virtual ~Scene() noexcept(false) { check(rtx_ctx_destroy(ctx_), "rtx_ctx_destroy"); }
    Scene(const Scene&) = delete;
    Scene& operator=(const Scene&) = delete;

    virtual unsigned int load() = 0;

    const Thing& operator[](unsigned int i) {
        check(rtx_thing_at(ctx_, i, &at_), "rtx_thing_at");
        return at_;
    }

    // create a mesh from the object's submeshes (concatenated)
    unsigned int add(Object& object) {
        std::vector<float3> v;
        std::vector<uint3> i;
        for (unsigned int o = 0; o < object.size(); o++) {
            float3* vces;
            unsigned int nv;
            uint3* ices;
            unsigned int ni;
            std::tie(vces, nv, ices, ni) = object[o];
            const uint32_t base = (uint32_t)v.size();
            v.insert(v.end(), vces, vces + nv);
            for (unsigned int k = 0; k < ni; k++) i.push_back(uint3{ices[k].x + base, ices[k].y + base, ices[k].z + base});
        }
        return add(Mesh(v.data(), (unsigned)v.size(), i.data(), (unsigned)i.size()));
    }
    unsigned int add(Mesh mesh) {
        uint32_t id = 0;
        check(rtx_mesh_add(ctx_, std::get<0>(mesh), std::get<1>(mesh), std::get<2>(mesh), std::get<3>(mesh), &id),
              "rtx_mesh_add");
        return id;
    }
    // instance of a mesh with identity transform; fills thing.vces/ices
    unsigned int add(Thing& thing, unsigned int object) {
        uint32_t id = 0;
        check(rtx_thing_add(ctx_, &thing, object, &id), "rtx_thing_add");
        return id;
    }
    bool set(unsigned int thing, const float* transform) { return rtx_thing_set(ctx_, thing, transform) == RTX_SUCCESS; }
    bool get(unsigned int thing, float* transform) { return rtx_thing_get(ctx_, thing, transform) == RTX_SUCCESS; }

    void build(TraversableHandle* is_handle, void* stream = nullptr) {
        check(rtx_scene_build(ctx_, stream, is_handle), "rtx_scene_build");
    }
    void update(TraversableHandle is_handle, void* stream = nullptr) {
        check(rtx_scene_update(ctx_, is_handle, stream), "rtx_scene_update");
    }
    // layout of the next build: RTX_BUILD_FLAT (default) or RTX_BUILD_TWO_LEVEL
    // (the reference's IAS over one GAS per mesh, optx/scene.cxx:38-273)
    void build_mode(int mode) { check(rtx_scene_set_build_mode(ctx_, mode), "rtx_scene_set_build_mode"); }
    int build_mode() {
        int m = 0;
        check(rtx_scene_get_build_mode(ctx_, &m), "rtx_scene_get_build_mode");
        return m;
    }
    rtx_ctx* context() { return ctx_; }

    // launch tuning (no reference counterpart); fields left 0 keep their value
    rtx_tuning tuning() {
        rtx_tuning t{};
        check(rtx_tuning_get(ctx_, &t), "rtx_tuning_get");
        return t;
    }
    void tune(const rtx_tuning& t) { check(rtx_tuning_set(ctx_, &t), "rtx_tuning_set"); }
    // material parameters, device float4 per thing (albedo rgb, fuzz)
    void set_materials(TraversableHandle h, const float* d_params, void* stream = nullptr) {
        check(rtx_scene_set_materials(ctx_, h, d_params, stream), "rtx_scene_set_materials");
    }
    void get_materials(TraversableHandle h, float* d_params, void* stream = nullptr) {
        check(rtx_scene_get_materials(ctx_, h, d_params, stream), "rtx_scene_get_materials");
    }

   private:
    rtx_ctx* ctx_ = nullptr;
    Thing at_{};
};

// Owns the per-pixel buffers and writes their pointers into the caller's
// launch parameters (optx/launcher.cxx:24-47).
class Launcher {
   public:
    Launcher(Scene& scene, LpGeneral& lp) : scene_(scene), lp_(lp) { resize(lp.image_w, lp.image_h); }
    ~Launcher() noexcept(false) { free(); }

    void resize(unsigned int w, unsigned int h) {
        free();
        const size_t n = (size_t)w * h;
        alloc((void**)&lp_.rawRGB, sizeof(float3) * n);
        alloc((void**)&lp_.rpp, sizeof(uint32_t) * n);
        alloc((void**)&lp_.normals, sizeof(float3) * n);
        alloc((void**)&lp_.albedos, sizeof(float3) * n);
        alloc((void**)&lp_.pick_id, sizeof(uint32_t));
        npix_ = n;
    }
    // synchronous, as the reference (optx/launcher.cxx:70)
    void ignite(void* stream = nullptr, bool once = false) {
        if (!once && (size_t)lp_.image_w * lp_.image_h > npix_)
            throw std::runtime_error("rtx error: image larger than the Launcher's buffers");
        check(rtx_launch(scene_.context(), &lp_, once ? 1 : 0, stream), "rtx_launch");
        check(rtx_synchronize(), "rtx_synchronize");
    }

    // dL/d material params (device float4 per thing) given dL/d rawRGB (device
    // float3 per pixel); rawRGB must hold the forward image of this lp
    void adjoint(const float3* d_dl_drgb, float* d_dparams, void* stream = nullptr) {
        check(rtx_launch_adjoint(scene_.context(), &lp_, d_dl_drgb, d_dparams, stream), "rtx_launch_adjoint");
        check(rtx_synchronize(), "rtx_synchronize");
    }

   private:
    void alloc(void** p, size_t bytes) {
        check(rtx_malloc(p, bytes), "rtx_malloc");
        owned_.push_back(*p);
    }
    void free() {
        for (void* p : owned_) rtx_free(p);
        owned_.clear();
        lp_.rawRGB = lp_.normals = lp_.albedos = nullptr;
        lp_.rpp = lp_.pick_id = nullptr;
        npix_ = 0;
    }
    Scene& scene_;
    LpGeneral& lp_;
    std::vector<void*> owned_;
    size_t npix_ = 0;
};

// Many frames in one persistent launch (no reference counterpart); frame i
// equals Launcher::ignite of lps[i].  The first launch on a stream captures a
// hipGraph that later launches replay.
class Batch {
   public:
    Batch(Scene& scene, const LpGeneral* lps, unsigned int n, bool use_graph = true) {
        check(rtx_batch_create(scene.context(), lps, n, use_graph ? 1 : 0, &b_), "rtx_batch_create");
    }
    ~Batch() noexcept(false) { check(rtx_batch_destroy(b_), "rtx_batch_destroy"); }
    Batch(const Batch&) = delete;
    Batch& operator=(const Batch&) = delete;
    void launch(void* stream, bool sync = true) {
        check(rtx_batch_launch(b_, stream), "rtx_batch_launch");
        if (sync) check(rtx_synchronize(), "rtx_synchronize");
    }

   private:
    rtx_batch* b_ = nullptr;
};

inline void pp_sRGB(const float3* src, uchar4* dst, int w, int h, void* stream = nullptr) {
    check(rtx_pp_srgb(src, dst, w, h, stream), "rtx_pp_srgb");
}
inline void pp_none(const float3* src, uchar4* dst, int w, int h, void* stream = nullptr) {
    check(rtx_pp_none(src, dst, w, h, stream), "rtx_pp_none");
}

}  // namespace rtx
