// Host-side sanitizer driver (test infrastructure): the library's host
// preparation (Scene::finalize rules, the parallel binned-SAH build, the
// quantized 4- and 8-wide trees, light and environment tables) and the device
// logic compiled for the CPU (traversal in both widths, exact resolution,
// shading, a full render), run over scene files written by
// tests/test_sanitizers.py.  Built with AddressSanitizer + UndefinedBehavior-
// Sanitizer, and with ThreadSanitizer for the threaded host build.
//
//   sanitize_check <scene.bin>...   exits 0 when every step succeeds
#include <cstdio>
#include <cstring>
#include <fstream>
#include <random>
#include <stdexcept>
#include <vector>

#include "../../include/crgpu.h"

extern "C" {
void* emu_scene_create(const crg_scene_desc* d);
void emu_scene_destroy(void* s);
int emu_scene_check(void* s);
void* emu_env_create(const crg_env_desc* d);
void emu_env_destroy(void* e);
int emu_trace(void* s, const crg_ray* rays, uint64_t n, crg_hit* hits, int any, uint64_t* fallbacks, uint64_t schedule);
int emu_render(void* s, void* e, const crg_camera* camera, const crg_options* o, float* film, crg_stats* stats);
const char* emu_last_error(void);
}

namespace {

template <class T>
void get(std::ifstream& f, T* p, size_t n) {
    f.read(reinterpret_cast<char*>(p), std::streamsize(n * sizeof(T)));
    if (!f) throw std::runtime_error("truncated scene file");
}

int run(const char* path) {
    std::ifstream f(path, std::ios::binary);
    uint64_t counts[3];
    get(f, counts, 3);
    const size_t n = counts[0];
    std::vector<float> pos(9 * n), nrm(9 * n);
    std::vector<uint32_t> mid(n);
    std::vector<crg_material> mats(counts[1]);
    std::vector<crg_light> lights(counts[2]);
    crg_camera cam{};
    float envc[3];
    crg_options opt{};
    get(f, pos.data(), pos.size());
    get(f, nrm.data(), nrm.size());
    get(f, mid.data(), n);
    get(f, mats.data(), mats.size());
    get(f, lights.data(), lights.size());
    get(f, &cam, 1);
    get(f, envc, 3);
    get(f, &opt, 1);
    crg_scene_desc d{};
    d.triangle_count = n;
    d.positions = pos.data();
    d.normals = nrm.data();
    d.material_ids = mid.data();
    d.material_count = uint32_t(mats.size());
    d.materials = mats.data();
    d.light_count = uint32_t(lights.size());
    d.lights = lights.data();
    void* s = emu_scene_create(&d);
    if (!s) {
        std::printf("%s: scene: %s\n", path, emu_last_error());
        return 1;
    }
    int rc = emu_scene_check(s);
    if (rc) std::printf("%s: tree check %d\n", path, rc);
    // An image environment (a small synthetic sky) exercises the table build.
    std::vector<float> sky(64 * 32 * 4);
    for (int y = 0; y < 32; ++y)
        for (int x = 0; x < 64; ++x)
            for (int c = 0; c < 4; ++c) sky[4 * (y * 64 + x) + c] = c == 3 ? 1.0f : 0.2f + float((x * 7 + y * 3 + c) % 11);
    crg_env_desc ed{};
    ed.width = 64;
    ed.height = 32;
    ed.rgba = sky.data();
    ed.rotation_degrees = 30;
    ed.intensity = 1.5f;
    void* e = emu_env_create(&ed);
    if (!e) {
        std::printf("%s: environment: %s\n", path, emu_last_error());
        emu_scene_destroy(s);
        return 1;
    }
    // Rays from around the camera in random directions, both tree widths.
    std::mt19937_64 g(7);
    std::uniform_real_distribution<float> u(-1.0f, 1.0f);
    std::vector<crg_ray> rays(2000);
    for (auto& r : rays) {
        r = crg_ray{};
        for (int k = 0; k < 3; ++k) {
            r.origin[k] = cam.position[k] + 0.1f * u(g);
            r.direction[k] = u(g);
        }
        r.tmax = 1e30f;
        r.ignore = 0xffffffffu;
    }
    std::vector<crg_hit> h4(rays.size()), h8(rays.size());
    uint64_t fb = 0;
    for (int any = 0; any < 2 && !rc; ++any) {
        rc |= emu_trace(s, rays.data(), rays.size(), h4.data(), any, &fb, 1);
        rc |= emu_trace(s, rays.data(), rays.size(), h8.data(), any, &fb, 9);
        // Closest hits agree completely; an any-hit query reports only whether the
        // segment is blocked (which blocker is met first depends on the tree).
        bool same = true;
        for (size_t i = 0; i < h4.size() && !rc; ++i)
            same = same && (any ? h4[i].hit == h8[i].hit : std::memcmp(&h4[i], &h8[i], sizeof(crg_hit)) == 0);
        if (!rc && !same) {
            std::printf("%s: 4- and 8-wide trees disagree (%s)\n", path, any ? "any hit" : "closest hit");
            rc = 1;
        }
    }
    std::vector<float> film(size_t(opt.width) * size_t(opt.height) * 3);
    crg_stats st{};
    if (!rc && emu_render(s, e, &cam, &opt, film.data(), &st)) {
        std::printf("%s: render: %s\n", path, emu_last_error());
        rc = 1;
    }
    emu_env_destroy(e);
    emu_scene_destroy(s);
    if (!rc) std::printf("%s: ok (%zu triangles, %llu samples)\n", path, n, (unsigned long long)st.camera_samples);
    return rc;
}

}  // namespace

int main(int argc, char** argv) {
    int rc = 0;
    for (int i = 1; i < argc; ++i) {
        try {
            rc |= run(argv[i]);
        } catch (const std::exception& ex) {
            std::printf("%s: %s\n", argv[i], ex.what());
            rc = 1;
        }
    }
    return rc;
}
