// TEST INFRASTRUCTURE ONLY — never linked into the product library.
//
// C-ABI probe harness around the reference renderer's own sources, which the
// oracle/Makefile compiles from where they lie (src/geometry.cpp, src/bsdf.cpp,
// src/render.cpp of the reference).  It turns crgpu.h descriptors into the
// reference's cr::Scene / cr::Environment / cr::Camera types and calls the
// reference functions unchanged.  Used to (1) write golden vectors under
// tests/golden/, (2) cross-check the restatement in cr_oracle.cpp, and
// (3) time the reference CPU renderer as bench.py's cpu_baseline.
#include "probe.h"
#include "render.h"
#include "../include/crgpu_cr.hpp"

#include <atomic>
#include <cmath>
#include <cstdio>
#include <cstring>
#include <filesystem>
#include <unistd.h>
#include <vector>
#include <exception>
#include <memory>
#include <string>

namespace {
thread_local std::string lastError;

cr::Vec3 v3(const float* p) { return {p[0], p[1], p[2]}; }
void put3(float* out, cr::Vec3 v) { out[0] = v.x; out[1] = v.y; out[2] = v.z; }

cr::TextureSlot slotOf(const crg_texture_slot& s) {
    cr::TextureSlot t;
    t.texture = s.texture; t.texcoord = s.texcoord;
    t.offset = {s.offset[0], s.offset[1]}; t.scale = {s.scale[0], s.scale[1]}; t.rotation = s.rotation;
    return t;
}

cr::Material materialOf(const crg_material& m) {
    cr::Material out;
    out.baseColor = {m.base_color[0], m.base_color[1], m.base_color[2], m.base_color[3]};
    out.metallic = m.metallic; out.roughness = m.roughness;
    out.emission = v3(m.emission);
    out.transmission = m.transmission; out.ior = m.ior;
    out.thickness = m.thickness; out.attenuationDistance = m.attenuation_distance;
    out.attenuationColor = v3(m.attenuation_color);
    out.doubleSided = m.double_sided != 0;
    out.alphaMode = m.alpha_mode ? cr::AlphaMode::Mask : cr::AlphaMode::Opaque;
    out.alphaCutoff = m.alpha_cutoff; out.normalScale = m.normal_scale;
    out.baseColorTexture = slotOf(m.base_color_tex);
    out.metallicRoughnessTexture = slotOf(m.metallic_roughness_tex);
    out.normalTexture = slotOf(m.normal_tex);
    out.emissiveTexture = slotOf(m.emissive_tex);
    out.transmissionTexture = slotOf(m.transmission_tex);
    return out;
}

cr::Camera cameraOf(const crg_camera& c) {
    cr::Camera out;
    out.position = v3(c.position); out.forward = v3(c.forward); out.right = v3(c.right); out.up = v3(c.up);
    out.yfov = c.yfov; out.aspect = c.aspect; out.nearClip = c.near_clip; out.farClip = c.far_clip;
    return out;
}

cr::Ray rayOf(const crg_ray& r) {
    cr::Ray out;
    out.origin = v3(r.origin); out.direction = v3(r.direction);
    out.tmin = r.tmin; out.tmax = r.tmax; out.originLow = v3(r.origin_low);
    return out;
}

void putRay(crg_ray* out, const cr::Ray& r) {
    put3(out->origin, r.origin); put3(out->direction, r.direction);
    out->tmin = r.tmin; out->tmax = r.tmax; put3(out->origin_low, r.originLow);
    out->ignore = 0xffffffffu;
}

cr::Surface surfaceOf(const probe_surface& s) {
    cr::Surface out;
    out.position = v3(s.position); out.positionLow = v3(s.position_low);
    out.geometricNormal = v3(s.geometric_normal); out.normal = v3(s.normal);
    out.tangent = v3(s.tangent); out.bitangent = v3(s.bitangent);
    out.baseColor = v3(s.base_color); out.emission = v3(s.emission); out.attenuationColor = v3(s.attenuation_color);
    out.metallic = s.metallic; out.roughness = s.roughness; out.transmission = s.transmission; out.ior = s.ior;
    out.thickness = s.thickness; out.attenuationDistance = s.attenuation_distance;
    out.material = s.material; out.frontFace = s.front_face != 0; out.doubleSided = s.double_sided != 0;
    return out;
}

void putSurface(probe_surface* out, const cr::Surface& s) {
    put3(out->position, s.position); put3(out->position_low, s.positionLow);
    put3(out->geometric_normal, s.geometricNormal); put3(out->normal, s.normal);
    put3(out->tangent, s.tangent); put3(out->bitangent, s.bitangent);
    put3(out->base_color, s.baseColor); put3(out->emission, s.emission); put3(out->attenuation_color, s.attenuationColor);
    out->metallic = s.metallic; out->roughness = s.roughness; out->transmission = s.transmission; out->ior = s.ior;
    out->thickness = s.thickness; out->attenuation_distance = s.attenuationDistance;
    out->material = s.material; out->front_face = s.frontFace; out->double_sided = s.doubleSided;
}

struct SceneBox {
    cr::Scene scene;
    cr::Bvh bvh;
};

template <class F>
int guarded(F&& f) {
    try {
        f();
        return 0;
    } catch (const std::bad_alloc&) {
        lastError = "out of memory";
        return CRG_ENOMEM;
    } catch (const std::exception& e) {
        lastError = e.what();
        return CRG_EINVAL;
    }
}
}  // namespace

extern "C" {

const char* ref_last_error(void) { return lastError.c_str(); }

uint64_t ref_sample_seed(uint64_t seed, uint64_t pixel, uint64_t sample) { return cr::sampleSeed(seed, pixel, sample); }

void ref_rng_draw(uint64_t* state, int count, float* uniforms) {
    cr::Rng rng(*state);
    for (int i = 0; i < count; ++i) uniforms[i] = rng.uniform();
    *state = rng.state;
}

int ref_camera_generate(const crg_camera* camera, float px, float py, int width, int height, crg_ray* out) {
    return guarded([&] { putRay(out, cameraOf(*camera).generate(px, py, width, height)); });
}

void* ref_scene_create(const crg_scene_desc* d) {
    auto box = std::make_unique<SceneBox>();
    const int rc = guarded([&] {
        cr::Scene& s = box->scene;
        for (uint32_t i = 0; i < d->material_count; ++i) s.materials.push_back(materialOf(d->materials[i]));
        for (uint32_t i = 0; i < d->image_count; ++i) {
            cr::Image im;
            im.width = d->images[i].width; im.height = d->images[i].height;
            const size_t n = static_cast<size_t>(im.width) * static_cast<size_t>(im.height);
            im.pixels.resize(n);
            for (size_t p = 0; p < n; ++p)
                im.pixels[p] = {d->images[i].rgba[4 * p], d->images[i].rgba[4 * p + 1], d->images[i].rgba[4 * p + 2],
                                d->images[i].rgba[4 * p + 3]};
            s.images.push_back(std::move(im));
        }
        for (uint32_t i = 0; i < d->texture_count; ++i) {
            cr::Texture t;
            t.image = d->textures[i].image;
            t.wrapS = static_cast<cr::Wrap>(d->textures[i].wrap_s);
            t.wrapT = static_cast<cr::Wrap>(d->textures[i].wrap_t);
            t.linear = d->textures[i].linear != 0;
            s.textures.push_back(t);
        }
        for (uint32_t i = 0; i < d->light_count; ++i) {
            const crg_light& l = d->lights[i];
            cr::PunctualLight p;
            p.type = static_cast<cr::LightType>(l.type);
            p.position = v3(l.position); p.direction = v3(l.direction); p.color = v3(l.color);
            p.intensity = l.intensity; p.range = l.range; p.innerCone = l.inner_cone; p.outerCone = l.outer_cone;
            s.lights.push_back(p);
        }
        s.triangles.resize(d->triangle_count);
        for (uint64_t t = 0; t < d->triangle_count; ++t) {
            cr::Triangle& tri = s.triangles[t];
            tri.material = d->material_ids[t];
            for (int k = 0; k < 3; ++k) {
                cr::Vertex& v = tri.v[k];
                v.position = v3(d->positions + 9 * t + 3 * k);
                if (d->normals) v.normal = v3(d->normals + 9 * t + 3 * k);
                if (d->tangents) {
                    const float* q = d->tangents + 12 * t + 4 * k;
                    v.tangent = {q[0], q[1], q[2], q[3]};
                }
                if (d->uvs) {
                    const float* q = d->uvs + 12 * t + 4 * k;
                    v.uv[0] = {q[0], q[1]};
                    v.uv[1] = {q[2], q[3]};
                }
            }
        }
        s.finalize();
        box->bvh.build(s);
    });
    if (rc) return nullptr;
    return box.release();
}

void ref_scene_destroy(void* scene) { delete static_cast<SceneBox*>(scene); }
uint64_t ref_scene_triangles(void* scene) { return static_cast<SceneBox*>(scene)->scene.triangles.size(); }
uint64_t ref_scene_nodes(void* scene) { return static_cast<SceneBox*>(scene)->bvh.nodeCount(); }

int ref_trace(void* scene, const crg_ray* rays, uint64_t count, crg_hit* hits, int any) {
    const SceneBox& b = *static_cast<SceneBox*>(scene);
    return guarded([&] {
        for (uint64_t i = 0; i < count; ++i) {
            cr::Hit h;
            const bool found = b.bvh.intersect(b.scene, rayOf(rays[i]), h, any != 0, rays[i].ignore);
            hits[i].hit = found;
            hits[i].triangle = found ? h.triangle : 0xffffffffu;
            hits[i].t = found ? h.t : 0;
            hits[i].u = found ? h.u : 0;
            hits[i].v = found ? h.v : 0;
        }
    });
}

int ref_offset_ray(const float* position, const float* normal, const float* direction, float max_distance, crg_ray* out) {
    return guarded([&] { putRay(out, cr::offsetRay(v3(position), v3(normal), v3(direction), max_distance)); });
}

int ref_surface_at(void* scene, const crg_ray* ray, const crg_hit* hit, probe_surface* out) {
    const SceneBox& b = *static_cast<SceneBox*>(scene);
    return guarded([&] {
        cr::Hit h;
        h.t = hit->t; h.u = hit->u; h.v = hit->v; h.triangle = hit->triangle;
        putSurface(out, cr::surfaceAt(b.scene, rayOf(*ray), h));
    });
}

void ref_bsdf_sample(const probe_surface* s, const float* wo, uint64_t* rng, probe_bsdf_sample* out) {
    cr::Rng r(*rng);
    const cr::BsdfSample b = cr::sampleBsdf(surfaceOf(*s), v3(wo), r);
    *rng = r.state;
    put3(out->direction, b.direction); put3(out->weight, b.weight);
    out->pdf = b.pdf; out->eta = b.eta; out->delta = b.delta; out->transmitted = b.transmitted;
    out->valid = static_cast<bool>(b);
}

void ref_bsdf_eval(const probe_surface* s, const float* wo, const float* wi, float* value_pdf) {
    const cr::BsdfEval e = cr::evaluateBsdf(surfaceOf(*s), v3(wo), v3(wi));
    put3(value_pdf, e.value);
    value_pdf[3] = e.pdf;
}

void* ref_env_create(const crg_env_desc* d) {
    cr::Environment* env = nullptr;
    guarded([&] {
        if (!d->rgba) {
            env = new cr::Environment(v3(d->constant));
        } else {
            cr::Image im;
            im.width = d->width; im.height = d->height;
            const size_t n = static_cast<size_t>(im.width) * static_cast<size_t>(im.height);
            im.pixels.resize(n);
            for (size_t p = 0; p < n; ++p) im.pixels[p] = {d->rgba[4 * p], d->rgba[4 * p + 1], d->rgba[4 * p + 2], d->rgba[4 * p + 3]};
            env = new cr::Environment(std::move(im), d->rotation_degrees, d->intensity);
        }
    });
    return env;
}

void ref_env_destroy(void* env) { delete static_cast<cr::Environment*>(env); }

void ref_env_sample(void* env, uint64_t* rng, float* direction, float* pdf) {
    cr::Rng r(*rng);
    put3(direction, static_cast<cr::Environment*>(env)->sample(r, *pdf));
    *rng = r.state;
}

float ref_env_pdf(void* env, const float* direction) { return static_cast<cr::Environment*>(env)->pdf(v3(direction)); }

void ref_env_radiance(void* env, const float* direction, float* rgb) {
    put3(rgb, static_cast<cr::Environment*>(env)->radiance(v3(direction)));
}

static int render_box(const SceneBox& b, void* env, const crg_camera* camera, const crg_options* o, float* film,
                      crg_stats* stats) {
    return guarded([&] {
        cr::RenderOptions opt;
        opt.width = o->width; opt.height = o->height; opt.samples = o->samples; opt.bounces = o->bounces;
        opt.threads = o->threads; opt.seed = o->seed;
        opt.nextEventEstimation = o->next_event_estimation != 0; opt.russianRoulette = o->russian_roulette != 0;
        const cr::RenderResult r = cr::render(b.scene, b.bvh, cameraOf(*camera), *static_cast<cr::Environment*>(env), opt);
        std::memcpy(film, r.pixels.data(), r.pixels.size() * sizeof(cr::Vec3));
        if (stats) {
            std::memset(stats, 0, sizeof(*stats));
            stats->seconds = r.seconds;
            stats->camera_samples = r.cameraSamples;
            stats->boundary_retries = r.boundaryRetries;
            stats->workers = r.workers;
        }
    });
}

int ref_render(void* scene, void* env, const crg_camera* camera, const crg_options* o, float* film, crg_stats* stats) {
    return render_box(*static_cast<SceneBox*>(scene), env, camera, o, film, stats);
}

// glTF import through the reference's loader; the descriptor is the adapter's
// flattening of the loaded (not yet finalized) scene.
namespace {
struct GltfBox {
    cr::Scene scene;
    cr::gpu::SceneStorage storage;
    crg_scene_desc desc{};
};
}  // namespace

void* ref_gltf_load(const char* path) {
    auto box = std::make_unique<GltfBox>();
    const int rc = guarded([&] {
        box->scene = cr::io::loadGltf(path);
        box->desc = cr::gpu::describe(box->scene, box->storage);
    });
    return rc ? nullptr : box.release();
}

void ref_gltf_destroy(void* g) { delete static_cast<GltfBox*>(g); }
const crg_scene_desc* ref_gltf_desc(void* g) { return &static_cast<GltfBox*>(g)->desc; }
int ref_gltf_camera_count(void* g) { return int(static_cast<GltfBox*>(g)->scene.cameras.size()); }

int ref_gltf_camera(void* g, int index, crg_camera* out) {
    const auto& cams = static_cast<GltfBox*>(g)->scene.cameras;
    if (index < 0 || size_t(index) >= cams.size()) return CRG_EINVAL;
    *out = cr::gpu::camera_of(cams[size_t(index)]);
    return CRG_OK;
}

int ref_gltf_render_loaded(void* g, void* env, const crg_camera* camera, const crg_options* o, float* film,
                           crg_stats* stats) {
    SceneBox b;
    const int rc = guarded([&] {
        b.scene = static_cast<GltfBox*>(g)->scene;
        b.scene.finalize();
        b.bvh.build(b.scene);
    });
    return rc ? rc : render_box(b, env, camera, o, film, stats);
}


// Film encode through the reference's own writers: writePng / writeHdr to a
// temporary file; the PNG is decoded again with the reference's decoder.
static std::filesystem::path film_temp(const char* ext) {
    static std::atomic<unsigned> counter{0};
    return std::filesystem::temp_directory_path() /
           ("crref_" + std::to_string(::getpid()) + "_" + std::to_string(counter++) + ext);
}
static std::vector<cr::Vec3> film_vec(const float* film, int width, int height) {
    std::vector<cr::Vec3> v(size_t(std::max(width, 0)) * size_t(std::max(height, 0)));
    for (size_t i = 0; i < v.size(); ++i) v[i] = {film[3 * i], film[3 * i + 1], film[3 * i + 2]};
    return v;
}
static std::vector<unsigned char> read_all(const std::filesystem::path& p) {
    std::FILE* f = std::fopen(p.c_str(), "rb");
    std::vector<unsigned char> b;
    if (!f) return b;
    unsigned char buf[65536];
    size_t n;
    while ((n = std::fread(buf, 1, sizeof(buf), f)) > 0) b.insert(b.end(), buf, buf + n);
    std::fclose(f);
    return b;
}

int ref_film_png8(const float* film, int width, int height, float exposure, uint8_t* rgb8) {
    return guarded([&] {
        const auto path = film_temp(".png");
        struct Rm { std::filesystem::path p; ~Rm() { std::error_code e; std::filesystem::remove(p, e); } } rm{path};
        cr::io::writePng(path, width, height, film_vec(film, width, height), exposure);
        const auto bytes = read_all(path);
        const cr::Image im = cr::io::decodeImage(bytes.data(), bytes.size(), "png", false);
        for (size_t i = 0; i < im.pixels.size(); ++i) {
            rgb8[3 * i] = uint8_t(std::lround(im.pixels[i].x * 255.0f));
            rgb8[3 * i + 1] = uint8_t(std::lround(im.pixels[i].y * 255.0f));
            rgb8[3 * i + 2] = uint8_t(std::lround(im.pixels[i].z * 255.0f));
        }
    });
}

int ref_film_hdr(const float* film, int width, int height, uint8_t* out, uint64_t capacity, uint64_t* size) {
    return guarded([&] {
        const auto path = film_temp(".hdr");
        struct Rm { std::filesystem::path p; ~Rm() { std::error_code e; std::filesystem::remove(p, e); } } rm{path};
        cr::io::writeHdr(path, width, height, film_vec(film, width, height));
        const auto bytes = read_all(path);
        *size = bytes.size();
        if (bytes.size() > capacity) throw std::runtime_error("HDR output buffer too small");
        std::memcpy(out, bytes.data(), bytes.size());
    });
}

int ref_film_write(const float* film, int width, int height, float exposure, const char* png_path,
                   const char* hdr_path) {
    return guarded([&] {
        const auto v = film_vec(film, width, height);
        cr::io::writePng(png_path, width, height, v, exposure);
        cr::io::writeHdr(hdr_path, width, height, v);
    });
}
}
