// rtwo.cpp -- batch renderer of the RTOW final scene over librtx.so, the
// batch path of optx/rtwo.cxx:78-620 with the OptiX pipeline replaced by the
// C-ABI (through include/rtx/rtx.hpp).
//
//   rtwo [-g WxH|NAME] [-a W:H] [-s spp] [-d depth] [-A RPP] [-G] [-D TYPE] [-S] [-q] [-v] [-t] [-h]
//        [-l level|mixed] [-r seed|glibc] [-I] > out.ppm
//
// The reference's batch options (optx/args.cxx:23-121) with their long
// forms: -g/--geometry takes WxH or a named resolution (optx/args.h:20-42),
// -a/--aspect-ratio W:H sets the height from the width, -s/--samples-per-pixel,
// -d/--trace-depth, -A/--print-aov RPP, -G/--print-guides, -S/--print-statistics,
// -q/--quiet/--silent, -h/--help/--usage, -v/--verbose and -t/--trace-sm
// (accepted; the latter only matters to the interactive viewer).
// -D/--apply-denoiser TYPE is accepted, but the OptiX denoiser is not part
// of this build: the image stays as rendered, and -G prints the guides
// (normals, albedos) only with -D, as the reference does.
// Options of this driver only: -l the subdivision level of every sphere
// (default: the reference's mixed levels), -r the scene seed (glibc: the
// reference's own rand() stream, the scene its batch binary renders), -I the
// reference's two-level layout (an instance structure over one structure
// per mesh) instead of the flat one.
//
// Output on stdout, top row first: the guides as PPM P3 (with -D -G,
// optx/rtwo.cxx:554-558, 639-654), the image as PPM P3
// (optx/rtwo.cxx:622-637), rays per pixel as PGM P2 with -A RPP
// (optx/rtwo.cxx:656-666); with -S the line "pixels rays milliseconds fps"
// on stderr (optx/rtwo.cxx:578-591).
#include <getopt.h>

#include <chrono>
#include <cmath>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <iostream>
#include <map>
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

// util::rnd() of the reference's host code (optx/util_cpu.h:90-91): the C
// library's rand() over (RAND_MAX + 1), from srand(1) -- the stream
// RTWO::load draws from, since it never seeds
struct GlibcRand {
    GlibcRand() { std::srand(1); }
    float operator()() { return (float)std::rand() / ((float)RAND_MAX + 1.f); }
    float range(float lo, float hi) { return lo + (*this)() * (hi - lo); }
};

// RTWO::load (optx/rtwo.cxx:133-245) drawn from one seeded Frand48 stream
// (per cell select, center x, center z, then the material draws, triples x
// first), or with -r glibc from the reference's own stream, whose V::rnd()
// triples g++ draws z first (it evaluates make_float3's arguments right to
// left): then the scene is the one the reference's batch binary renders.
class RTWO : public rtx::Scene {
   public:
    RTWO(uint32_t seed, int level, bool glibc = false) : seed_(seed), level_(level), glibc_(glibc) {}

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
        if (glibc_) {
            GlibcRand rnd;
            return recipe(gas, rnd, true);
        }
        Frand48 rnd(seed_);
        return recipe(gas, rnd, false);
    }

   private:
    // one triple of draws (V::rnd, optx/v.h:80-81) in the stream's order
    template <class R>
    static void triple(R& rnd, bool zfirst, float lo, float hi, float& x, float& y, float& z) {
        if (zfirst) {
            z = rnd.range(lo, hi);
            y = rnd.range(lo, hi);
            x = rnd.range(lo, hi);
        } else {
            x = rnd.range(lo, hi);
            y = rnd.range(lo, hi);
            z = rnd.range(lo, hi);
        }
    }
    template <class R>
    unsigned int recipe(const unsigned int (&gas)[10], R& rnd, bool zfirst) {
        auto pick = [&](int mixed) { return gas[level_ >= 0 ? level_ : mixed]; };
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
                        float r1, g1, b1, r2, g2, b2;
                        triple(rnd, zfirst, 0.f, 1.f, r1, g1, b1);
                        triple(rnd, zfirst, 0.f, 1.f, r2, g2, b2);
                        thing.optics.type = RTX_OPTICS_DIFFUSE;
                        thing.optics.diffuse.albedo = {r1 * r2, g1 * g2, b1 * b2};
                        place(add(thing, pick(6)), .2f, cx, .2f, cz);
                    } else if (select < .95f) {
                        float r1, g1, b1;
                        triple(rnd, zfirst, .5f, 1.f, r1, g1, b1);
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

    uint32_t seed_;
    int level_;
    bool glibc_;
};

void usage() {
    std::cerr << "usage: rtwo [-g WxH|NAME] [-a W:H] [-s spp] [-d depth] [-A RPP] [-G] [-D TYPE] [-S] [-q] [-v] [-t] [-h]\n"
                 "            [-l level|mixed] [-r seed|glibc] [-I]\n";
}

// named resolutions of the -g option (optx/args.h:20-42)
const std::map<std::string, std::pair<int, int>> kResolutions = {
    {"CGA", {320, 200}},     {"HVGA", {480, 320}},    {"VGA", {640, 480}},     {"WVGA", {800, 480}},
    {"SVGA", {800, 600}},    {"XGA", {1024, 768}},    {"HD", {1280, 720}},     {"SXGA", {1280, 1024}},
    {"UXGA", {1600, 1200}},  {"FULLHD", {1920, 1080}}, {"2K", {2048, 1080}},   {"QXGA", {2048, 1536}},
    {"UWHD", {2560, 1080}},  {"WQHD", {2560, 1440}},  {"WQXGA", {2560, 1600}}, {"UWQHD", {3440, 1440}},
    {"UHD-1", {3840, 2160}}, {"4K", {4096, 2160}},    {"5K-UW", {5120, 2160}}, {"5K", {5120, 2880}},
    {"UHD-2", {7680, 4320}}};

// float3 image as PPM P3, top row first, components clamped to [0, 1] and
// scaled by 255 with truncation (optx/rtwo.cxx:639-654)
void print_float3(const std::vector<rtx::float3>& img, unsigned w, unsigned h) {
    std::cout << "P3\n" << w << ' ' << h << '\n' << 255 << '\n';
    auto c = [](float v) { return (int)((v < 0.f ? 0.f : v > 1.f ? 1.f : v) * 255); };
    for (int y = (int)h - 1; y >= 0; --y)
        for (unsigned x = 0; x < w; ++x) {
            const rtx::float3 p = img[(size_t)w * y + x];
            std::cout << c(p.x) << ' ' << c(p.y) << ' ' << c(p.z) << '\n';
        }
    std::cout << std::endl;
}

}  // namespace

int main(int argc, char* argv[]) {
    // optx/rtwo.cxx:89-98: 1280x720, 50 spp, depth 16 unless given
    int gw = -1, gh = -1;
    unsigned spp = 50, depth = 16;
    int level = -1;
    uint32_t seed = 0;
    bool glibc = false;
    bool stats = false, quiet = false, aov_rpp = false, two_level = false, guides = false, denoise = false;
    int flag_v = 0, flag_t = 0, flag_q = 0, flag_h = 0, flag_G = 0, flag_S = 0;
    const struct option lopts[] = {
        {"geometry", required_argument, nullptr, 'g'},     {"aspect-ratio", required_argument, nullptr, 'a'},
        {"samples-per-pixel", required_argument, nullptr, 's'}, {"trace-depth", required_argument, nullptr, 'd'},
        {"verbose", no_argument, &flag_v, 1},              {"trace-sm", no_argument, &flag_t, 1},
        {"quiet", no_argument, &flag_q, 1},                {"silent", no_argument, &flag_q, 1},
        {"help", no_argument, &flag_h, 1},                 {"usage", no_argument, &flag_h, 1},
        {"print-aov", required_argument, nullptr, 'A'},    {"print-guides", no_argument, &flag_G, 1},
        {"print-statistics", no_argument, &flag_S, 1},     {"apply-denoiser", required_argument, nullptr, 'D'},
        {nullptr, 0, nullptr, 0}};
    int c;
    while ((c = getopt_long(argc, argv, "g:a:s:d:vtqhA:GSD:l:r:I", lopts, nullptr)) != -1) {
        switch (c) {
            case 0: break;   // a flag set through lopts
            case 'g': {
                const auto it = kResolutions.find(optarg);
                if (it != kResolutions.end()) {
                    gw = it->second.first;
                    gh = it->second.second;
                } else if (std::sscanf(optarg, "%dx%d", &gw, &gh) != 2 || gw < 1 || gh < 1) {
                    gw = gh = -1;   // optx/args.cxx:30-34: a bad geometry keeps the default
                }
                break;
            }
            case 'a': {
                float aw = 0.f, ah = 0.f;   // optx/args.cxx:40-45
                if (std::sscanf(optarg, "%f:%f", &aw, &ah) == 2 && gw > 0 && aw > 0.f && ah > 0.f)
                    gh = (int)((float)gw * ah / aw + .5f);
                break;
            }
            case 's': spp = (unsigned)std::abs(std::atoi(optarg)); break;
            case 'd': depth = (unsigned)std::abs(std::atoi(optarg)); break;
            case 'v': flag_v = 1; break;
            case 't': flag_t = 1; break;
            case 'q': flag_q = 1; break;
            case 'h': flag_h = 1; break;
            case 'G': flag_G = 1; break;
            case 'S': flag_S = 1; break;
            case 'A': {
                // comma-separated AOV names (optx/args.cxx:64-80)
                std::string list = optarg;
                size_t pos = 0;
                while (pos <= list.size()) {
                    const size_t e = list.find(',', pos);
                    const std::string aov = list.substr(pos, e == std::string::npos ? std::string::npos : e - pos);
                    if (aov == "RPP") aov_rpp = true;
                    else std::cerr << "rtwo: unknown argument for option A ignored -- " << aov << std::endl;
                    if (e == std::string::npos) break;
                    pos = e + 1;
                }
                break;
            }
            case 'D': {
                const std::string t = optarg;
                if (t == "SMP" || t == "NRM" || t == "ALB" || t == "NAA" || t == "AOV") {
                    denoise = true;
                    std::cerr << "rtwo: the OptiX denoiser is not part of this build; the image is not denoised"
                              << std::endl;
                } else {
                    std::cerr << "rtwo: unknown argument for option D ignored -- " << t << std::endl;
                }
                break;
            }
            case 'l': level = std::string(optarg) == "mixed" ? -1 : std::atoi(optarg); break;
            case 'r':
                if (std::strcmp(optarg, "glibc") == 0) glibc = true;
                else seed = (uint32_t)std::strtoul(optarg, nullptr, 10);
                break;
            case 'I': two_level = true; break;
            default: usage(); return 1;   // optx/args.cxx:96-98
        }
    }
    if (flag_h) { usage(); return 0; }
    (void)flag_v;
    (void)flag_t;
    quiet = flag_q != 0;
    stats = flag_S != 0;
    guides = flag_G != 0 && denoise;
    const unsigned w = gw > 0 ? (unsigned)gw : 1280u, h = gh > 0 ? (unsigned)gh : 720u;
    try {
        rtx::LpGeneral lp{};
        lp.image_w = w;
        lp.image_h = h;
        lp.spp = spp;
        lp.depth = depth;
        rtx::Camera camera;
        camera.set({13.f, 2.f, 3.f}, {0.f, 0.f, 0.f}, {0.f, 1.f, 0.f}, 20.f, (float)w / (float)h, .1f, 10.f);
        lp.camera = camera.pod();

        RTWO rtwo(seed, level, glibc);
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

        if (!quiet && guides) {
            std::vector<rtx::float3> g(n);
            rtx::check(rtx_memcpy_dtoh(g.data(), lp.normals, sizeof(rtx::float3) * n), "rtx_memcpy_dtoh");
            print_float3(g, w, h);
            rtx::check(rtx_memcpy_dtoh(g.data(), lp.albedos, sizeof(rtx::float3) * n), "rtx_memcpy_dtoh");
            print_float3(g, w, h);
        }
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
