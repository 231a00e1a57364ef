// rtwo.cpp -- batch renderer of the RTOW final scene over librtx.so, the
// batch path of optx/rtwo.cxx:78-620 with the OptiX pipeline replaced by the
// C-ABI (through include/rtx/rtx.hpp).
//
//   rtwo [-g WxH] [-s spp] [-d depth] [-l level|mixed] [-r seed] [-I] [-S] [-q] [-A RPP] > out.ppm
//
// -I builds the reference's two-level layout (an instance structure over
// one structure per mesh) instead of the flat one.
//
// Output: PPM P3 on stdout, top row first (optx/rtwo.cxx:622-637); with -A RPP
// also a PGM P2 of rays per pixel (optx/rtwo.cxx:656-666); with -S the line
// "pixels rays milliseconds fps" on stderr (optx/rtwo.cxx:578-591).
#include <chrono>
#include <cmath>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <iostream>
#include <string>
#include <vector>

#include "rtx/rtx.hpp"

namespace {

// Frand48 on the host (optx/frand48.h:17-35), for the scene recipe
struct Frand48 {
    uint64_t state;
    explicit Frand48(uint32_t seed) : state(seed) {
        for (int i = 0; i < 16; i++) (*this)();
    }
    float operator()() {
        state = 0x5DEECE66Dull * state + 0xBull;
        return float(state & 0xFFFFFFFFFFFFull) / float(0xFFFFFFFFFFFFull + 1ull);
    }
    float range(float lo, float hi) { return lo + (*this)() * (hi - lo); }
};

// RTWO::load (optx/rtwo.cxx:133-245) drawn from one seeded Frand48 stream:
// per cell select, center x, center z, then the material draws.
class RTWO : public rtx::Scene {
   public:
    RTWO(uint32_t seed, int level) : seed_(seed), level_(level) {}

    unsigned int load() override {
        // level_ < 0: the reference's mixed levels (optx/rtwo.cxx:138-145)
        unsigned int gas[10] = {};
        for (int l = 0; l <= 9; l++) {
            const bool need = level_ >= 0 ? l == level_ : (l == 3 || l == 6 || l == 8 || l == 9);
            if (need) {
                rtx::Sphere s(1.f, (unsigned)l);
                gas[l] = add(s.mesh());
            }
        }
        auto pick = [&](int mixed) { return gas[level_ >= 0 ? level_ : mixed]; };
        Frand48 rnd(seed_);
        rtx::Thing thing{};
        float xf[12] = {1, 0, 0, 0, 0, 1, 0, 0, 0, 0, 1, 0};
        auto place = [&](unsigned id, float r, float cx, float cy, float cz) {
            xf[0] = r; xf[5] = r; xf[10] = r;
            xf[3] = cx; xf[7] = cy; xf[11] = cz;
            set(id, xf);
        };
        thing.optics.type = RTX_OPTICS_DIFFUSE;
        thing.optics.diffuse.albedo = {.5f, .5f, .5f};
        place(add(thing, pick(9)), 1000.f, 0.f, -1000.f, 0.f);
        for (int a = -11; a < 11; a++) {
            for (int b = -11; b < 11; b++) {
                const float select = rnd();
                const float cx = (float)a + .9f * rnd();
                const float cz = (float)b + .9f * rnd();
                const float dx = cx - 4.f, dy = .2f - .2f, dz = cz - 0.f;
                if (std::sqrt(dx * dx + dy * dy + dz * dz) > .9f) {
                    if (select < .8f) {
                        const float r1 = rnd(), g1 = rnd(), b1 = rnd();
                        const float r2 = rnd(), g2 = rnd(), b2 = rnd();
                        thing.optics.type = RTX_OPTICS_DIFFUSE;
                        thing.optics.diffuse.albedo = {r1 * r2, g1 * g2, b1 * b2};
                        place(add(thing, pick(6)), .2f, cx, .2f, cz);
                    } else if (select < .95f) {
                        const float r1 = rnd.range(.5f, 1.f), g1 = rnd.range(.5f, 1.f), b1 = rnd.range(.5f, 1.f);
                        const float fz = rnd.range(0.f, .5f);
                        thing.optics.type = RTX_OPTICS_REFLECT;
                        thing.optics.reflect.albedo = {r1, g1, b1};
                        thing.optics.reflect.fuzz = fz;
                        place(add(thing, pick(6)), .2f, cx, .2f, cz);
                    } else {
                        thing.optics.type = RTX_OPTICS_REFRACT;
                        thing.optics.refract.index = 1.5f;
                        place(add(thing, pick(3)), .2f, cx, .2f, cz);
                    }
                }
            }
        }
        thing.optics.type = RTX_OPTICS_REFRACT;
        thing.optics.refract.index = 1.5f;
        place(add(thing, pick(8)), 1.f, 0.f, 1.f, 0.f);
        thing.optics.type = RTX_OPTICS_DIFFUSE;
        thing.optics.diffuse.albedo = {.4f, .2f, .1f};
        place(add(thing, pick(6)), 1.f, -4.f, 1.f, 0.f);
        thing.optics.type = RTX_OPTICS_REFLECT;
        thing.optics.reflect.albedo = {.7f, .6f, .5f};
        thing.optics.reflect.fuzz = 0.f;
        const unsigned id = add(thing, pick(3));
        place(id, 1.f, 4.f, 1.f, 0.f);
        return id + 1;
    }

   private:
    uint32_t seed_;
    int level_;
};

void usage() {
    std::cerr << "usage: rtwo [-g WxH] [-s spp] [-d depth] [-l level|mixed] [-r seed] [-I] [-S] [-q] [-A RPP]\n";
}

}  // namespace

int main(int argc, char* argv[]) {
    unsigned w = 1280, h = 720, spp = 50, depth = 16;
    int level = -1;
    uint32_t seed = 0;
    bool stats = false, quiet = false, aov_rpp = false, two_level = false;
    for (int i = 1; i < argc; i++) {
        const std::string a = argv[i];
        auto val = [&]() -> std::string {
            if (i + 1 >= argc) { usage(); std::exit(1); }
            return argv[++i];
        };
        if (a == "-g") {
            const std::string g = val();
            if (std::sscanf(g.c_str(), "%ux%u", &w, &h) != 2 || !w || !h) { usage(); return 1; }
        } else if (a == "-s") spp = (unsigned)std::atoi(val().c_str());
        else if (a == "-d") depth = (unsigned)std::atoi(val().c_str());
        else if (a == "-l") { const std::string l = val(); level = l == "mixed" ? -1 : std::atoi(l.c_str()); }
        else if (a == "-r") seed = (uint32_t)std::strtoul(val().c_str(), nullptr, 10);
        else if (a == "-S") stats = true;
        else if (a == "-I") two_level = true;
        else if (a == "-q") quiet = true;
        else if (a == "-A") aov_rpp = val() == "RPP";
        else if (a == "-h") { usage(); return 0; }
        else { usage(); return 1; }
    }
    try {
        rtx::LpGeneral lp{};
        lp.image_w = w;
        lp.image_h = h;
        lp.spp = spp;
        lp.depth = depth;
        rtx::Camera camera;
        camera.set({13.f, 2.f, 3.f}, {0.f, 0.f, 0.f}, {0.f, 1.f, 0.f}, 20.f, (float)w / (float)h, .1f, 10.f);
        lp.camera = camera.pod();

        RTWO rtwo(seed, level);
        if (two_level) rtwo.build_mode(RTX_BUILD_TWO_LEVEL);
        rtwo.load();
        rtwo.build(&lp.is_handle);

        rtx::Launcher launcher(rtwo, lp);
        const auto t0 = std::chrono::high_resolution_clock::now();
        launcher.ignite();
        const auto t1 = std::chrono::high_resolution_clock::now();

        const size_t n = (size_t)w * h;
        void* image = nullptr;
        rtx::check(rtx_malloc(&image, sizeof(rtx::uchar4) * n), "rtx_malloc");
        rtx::pp_sRGB(lp.rawRGB, (rtx::uchar4*)image, (int)w, (int)h);
        std::vector<rtx::uchar4> img(n);
        rtx::check(rtx_memcpy_dtoh(img.data(), image, sizeof(rtx::uchar4) * n), "rtx_memcpy_dtoh");
        rtx_free(image);
        std::vector<uint32_t> rpp(n);
        rtx::check(rtx_memcpy_dtoh(rpp.data(), lp.rpp, sizeof(uint32_t) * n), "rtx_memcpy_dtoh");

        if (!quiet) {
            std::cout << "P3\n" << w << ' ' << h << '\n' << 255 << '\n';
            for (int y = (int)h - 1; y >= 0; --y)
                for (unsigned x = 0; x < w; ++x) {
                    const rtx::uchar4 p = img[(size_t)w * y + x];
                    std::cout << (int)p.x << ' ' << (int)p.y << ' ' << (int)p.z << '\n';
                }
            std::cout << std::endl;
            if (aov_rpp) {
                std::cout << "P2\n" << w << ' ' << h << '\n' << 65535 << '\n';
                for (int y = (int)h - 1; y >= 0; --y)
                    for (unsigned x = 0; x < w; ++x) std::cout << rpp[(size_t)w * y + x] << '\n';
                std::cout << std::endl;
            }
        }
        if (stats) {
            const long long dt = std::chrono::duration_cast<std::chrono::milliseconds>(t1 - t0).count();
            unsigned long long sr = 0;
            for (uint32_t c : rpp) sr += c;
            std::fprintf(stderr, "%9u %12llu %4lld (pixel, rays, milliseconds) %6.2f fps\n", w * h, sr, dt,
                         dt > 0 ? 1000.f / dt : 0.f);
        }
    } catch (const std::exception& e) {
        std::cerr << "exception: " << e.what() << "\n";
        return 1;
    }
    return 0;
}
