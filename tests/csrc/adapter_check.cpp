// The reference-side drop-in, instantiated with the reference renderer's own
// types: cr::gpu::render with exactly the parameter list of cr::render
// (include/crgpu_render.hpp; reference src/render.h:53).
//
//   adapter_check <scene.bin> <film.out>   render the scene file on device 0
//   adapter_check --compile-only           exits after instantiation (CPU check)
//
// The scene file (written by tests/test_gpu_adapter.py) holds the C records of
// a scene with a constant environment: counts, positions, normals, material
// ids, crg_material / crg_light records, a crg_camera, the environment
// radiance and crg_options.  The program rebuilds a cr::Scene from them, renders
// it through the drop-in, and writes the film followed by a line per progress
// call.  It also checks the drop-in's two refusals: an image Environment
// (texels private) and options the reference rejects (its message).
#include <cstdio>
#include <cstring>
#include <fstream>
#include <iostream>
#include <vector>

#include "crgpu_render.hpp"

namespace {

template <class T>
void get(std::ifstream& f, T* p, size_t n) {
    f.read(reinterpret_cast<char*>(p), std::streamsize(n * sizeof(T)));
    if (!f) throw std::runtime_error("truncated scene file");
}

cr::Vec3 v3(const float* p) { return cr::Vec3{p[0], p[1], p[2]}; }

cr::TextureSlot slot(const crg_texture_slot& s) {
    cr::TextureSlot t;
    t.texture = s.texture;
    t.texcoord = s.texcoord;
    t.offset = cr::Vec2{s.offset[0], s.offset[1]};
    t.scale = cr::Vec2{s.scale[0], s.scale[1]};
    t.rotation = s.rotation;
    return t;
}

}  // namespace

int main(int argc, char** argv) {
    if (argc == 2 && std::strcmp(argv[1], "--compile-only") == 0) {
        std::printf("instantiated\n");
        return 0;
    }
    if (argc != 3) {
        std::fprintf(stderr, "usage: adapter_check <scene.bin> <film.out>\n");
        return 2;
    }
    std::ifstream f(argv[1], std::ios::binary);
    uint64_t counts[3];  // triangles, materials, lights
    get(f, counts, 3);
    const size_t n = counts[0];
    std::vector<float> pos(9 * n), nrm(9 * n);
    std::vector<uint32_t> mid(n);
    std::vector<crg_material> mats(counts[1]);
    std::vector<crg_light> lights(counts[2]);
    crg_camera cam{};
    float env[3];
    crg_options opt{};
    get(f, pos.data(), pos.size());
    get(f, nrm.data(), nrm.size());
    get(f, mid.data(), n);
    get(f, mats.data(), mats.size());
    get(f, lights.data(), lights.size());
    get(f, &cam, 1);
    get(f, env, 3);
    get(f, &opt, 1);

    cr::Scene scene;
    for (const crg_material& c : mats) {
        cr::Material m;
        m.baseColor = cr::Vec4{c.base_color[0], c.base_color[1], c.base_color[2], c.base_color[3]};
        m.metallic = c.metallic;
        m.roughness = c.roughness;
        m.emission = v3(c.emission);
        m.transmission = c.transmission;
        m.ior = c.ior;
        m.thickness = c.thickness;
        m.attenuationDistance = c.attenuation_distance;
        m.attenuationColor = v3(c.attenuation_color);
        m.doubleSided = c.double_sided != 0;
        m.alphaMode = c.alpha_mode ? cr::AlphaMode::Mask : cr::AlphaMode::Opaque;
        m.alphaCutoff = c.alpha_cutoff;
        m.normalScale = c.normal_scale;
        m.baseColorTexture = slot(c.base_color_tex);
        m.metallicRoughnessTexture = slot(c.metallic_roughness_tex);
        m.normalTexture = slot(c.normal_tex);
        m.emissiveTexture = slot(c.emissive_tex);
        m.transmissionTexture = slot(c.transmission_tex);
        scene.materials.push_back(m);
    }
    for (const crg_light& c : lights) {
        cr::PunctualLight l;
        l.type = static_cast<cr::LightType>(c.type);
        l.position = v3(c.position);
        l.direction = v3(c.direction);
        l.color = v3(c.color);
        l.intensity = c.intensity;
        l.range = c.range;
        l.innerCone = c.inner_cone;
        l.outerCone = c.outer_cone;
        scene.lights.push_back(l);
    }
    for (size_t t = 0; t < n; ++t) {
        cr::Triangle tri;
        tri.material = mid[t];
        for (int k = 0; k < 3; ++k) {
            tri.v[k].position = v3(&pos[9 * t + 3 * k]);
            tri.v[k].normal = v3(&nrm[9 * t + 3 * k]);
        }
        scene.triangles.push_back(tri);
    }
    scene.finalize();
    cr::Bvh bvh;
    bvh.build(scene);
    cr::Camera camera;
    camera.position = v3(cam.position);
    camera.forward = v3(cam.forward);
    camera.right = v3(cam.right);
    camera.up = v3(cam.up);
    camera.yfov = cam.yfov;
    camera.aspect = cam.aspect;
    camera.nearClip = cam.near_clip;
    camera.farClip = cam.far_clip;
    cr::RenderOptions options;
    options.width = opt.width;
    options.height = opt.height;
    options.samples = opt.samples;
    options.bounces = opt.bounces;
    options.threads = opt.threads;
    options.seed = opt.seed;
    options.nextEventEstimation = opt.next_event_estimation != 0;
    options.russianRoulette = opt.russian_roulette != 0;
    const cr::Environment environment(v3(env));

    std::vector<std::string> calls;
    const cr::Progress progress = [&](uint64_t done, uint64_t total, double seconds) {
        char line[128];
        std::snprintf(line, sizeof(line), "progress %llu %llu %.3f", (unsigned long long)done,
                      (unsigned long long)total, seconds);
        calls.push_back(line);
    };
    cr::RenderResult r;
    try {
        r = cr::gpu::render(scene, bvh, camera, environment, options, progress);
    } catch (const std::exception& e) {
        std::printf("render failed: %s\n", e.what());
        return 1;
    }
    std::ofstream out(argv[2], std::ios::binary);
    out.write(reinterpret_cast<const char*>(r.pixels.data()), std::streamsize(r.pixels.size() * sizeof(cr::Vec3)));
    std::printf("rendered %zu px, %llu samples, %llu replays, workers %d, %.3f s\n", r.pixels.size(),
                (unsigned long long)r.cameraSamples, (unsigned long long)r.boundaryRetries, r.workers, r.seconds);
    for (const auto& c : calls) std::printf("%s\n", c.c_str());

    // Refusals: an image environment, and options outside the reference's ranges.
    cr::Image sky;
    sky.width = 4;
    sky.height = 2;
    sky.pixels.assign(8, cr::Vec4{1, 1, 1, 1});
    try {
        cr::gpu::render(scene, bvh, camera, cr::Environment(sky, 0, 1), options);
        std::printf("image environment: accepted\n");
    } catch (const std::invalid_argument& e) {
        std::printf("image environment: %s\n", e.what());
    }
    cr::RenderOptions bad = options;
    bad.samples = 0;
    try {
        cr::gpu::render(scene, bvh, camera, environment, bad);
        std::printf("bad options: accepted\n");
    } catch (const std::runtime_error& e) {
        std::printf("bad options: %s\n", e.what());
    }
    return 0;
}
