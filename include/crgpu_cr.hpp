// crgpu_cr.hpp — header-only C++20 adapter from the reference renderer's types
// to the crgpu C ABI.  A maintainer of the reference adds this header and links
// libcrgpu.so; `cr::gpu::render(...)` then replaces `cr::render(...)` with the
// same result type (reference src/render.h:53): with a Device first and the
// reference's own arguments after it, or, through include/crgpu_render.hpp,
// with exactly the reference's parameter list.  `ResidentScene::intersect`
// replaces `Bvh::intersect` (reference src/scene.h:126-127), one ray or a batch.
//
// The adapter is a template over the reference types so this header compiles
// on its own; it is instantiated where the reference's render.h is included.
// A constant Environment converts directly.  An image Environment keeps its
// texels private, so for image lighting the caller passes a crg_env_desc made
// from the same HDR image, rotation and intensity it gave the Environment
// constructor (reference src/main.cpp:212).
#pragma once
#include <chrono>
#include <cstring>
#include <new>
#include <stdexcept>
#include <string>
#include <vector>

#include "crgpu.h"

namespace cr::gpu {

class Device {
public:
    explicit Device(int device = 0) {
        if (crg_context_create(device, &ctx_) != CRG_OK)
            throw std::runtime_error("no usable gfx950 device " + std::to_string(device));
    }
    ~Device() { crg_context_destroy(ctx_); }
    Device(const Device&) = delete;
    Device& operator=(const Device&) = delete;
    crg_ctx* handle() const { return ctx_; }
    // Map a crgpu status to the reference's exception types.
    void check(int rc) const {
        if (rc != CRG_OK) raise(rc, crg_last_error(ctx_));
    }
    // The same for a status whose text was read before later calls replaced it.
    [[noreturn]] static void raise(int rc, const std::string& message) {
        if (rc == CRG_ENOMEM) throw std::bad_alloc();
        throw std::runtime_error(message);
    }

private:
    crg_ctx* ctx_ = nullptr;
};

// Flatten a finalized-or-not reference Scene into the C descriptor.  The
// storage vectors own the flattened arrays for the descriptor's lifetime.
struct SceneStorage {
    std::vector<float> positions, normals, tangents, uvs;
    std::vector<uint32_t> material_ids;
    std::vector<crg_material> materials;
    std::vector<crg_light> lights;
    std::vector<crg_image> images;
    std::vector<std::vector<float>> pixels;
    std::vector<crg_texture> textures;
};

template <class Slot>
crg_texture_slot slot_of(const Slot& s) {
    return crg_texture_slot{s.texture, s.texcoord, {s.offset.x, s.offset.y}, {s.scale.x, s.scale.y}, s.rotation};
}

template <class Scene>
crg_scene_desc describe(const Scene& scene, SceneStorage& st) {
    const size_t n = scene.triangles.size();
    st.positions.resize(9 * n);
    st.normals.resize(9 * n);
    st.tangents.resize(12 * n);
    st.uvs.resize(12 * n);
    st.material_ids.resize(n);
    for (size_t t = 0; t < n; ++t) {
        const auto& tri = scene.triangles[t];
        st.material_ids[t] = tri.material;
        for (int k = 0; k < 3; ++k) {
            const auto& v = tri.v[k];
            const float p[3] = {v.position.x, v.position.y, v.position.z};
            const float nn[3] = {v.normal.x, v.normal.y, v.normal.z};
            std::memcpy(&st.positions[9 * t + 3 * k], p, sizeof(p));
            std::memcpy(&st.normals[9 * t + 3 * k], nn, sizeof(nn));
            const float tg[4] = {v.tangent.x, v.tangent.y, v.tangent.z, v.tangent.w};
            std::memcpy(&st.tangents[12 * t + 4 * k], tg, sizeof(tg));
            const float uv[4] = {v.uv[0].x, v.uv[0].y, v.uv[1].x, v.uv[1].y};
            std::memcpy(&st.uvs[12 * t + 4 * k], uv, sizeof(uv));
        }
    }
    for (const auto& m : scene.materials) {
        crg_material c{};
        c.base_color[0] = m.baseColor.x; c.base_color[1] = m.baseColor.y;
        c.base_color[2] = m.baseColor.z; c.base_color[3] = m.baseColor.w;
        c.metallic = m.metallic;
        c.roughness = m.roughness;
        c.emission[0] = m.emission.x; c.emission[1] = m.emission.y; c.emission[2] = m.emission.z;
        c.transmission = m.transmission;
        c.ior = m.ior;
        c.thickness = m.thickness;
        c.attenuation_distance = m.attenuationDistance;
        c.attenuation_color[0] = m.attenuationColor.x;
        c.attenuation_color[1] = m.attenuationColor.y;
        c.attenuation_color[2] = m.attenuationColor.z;
        c.double_sided = m.doubleSided;
        c.alpha_mode = static_cast<int32_t>(m.alphaMode);
        c.alpha_cutoff = m.alphaCutoff;
        c.normal_scale = m.normalScale;
        c.base_color_tex = slot_of(m.baseColorTexture);
        c.metallic_roughness_tex = slot_of(m.metallicRoughnessTexture);
        c.normal_tex = slot_of(m.normalTexture);
        c.emissive_tex = slot_of(m.emissiveTexture);
        c.transmission_tex = slot_of(m.transmissionTexture);
        st.materials.push_back(c);
    }
    for (const auto& l : scene.lights) {
        crg_light c{};
        c.type = static_cast<int32_t>(l.type);
        c.position[0] = l.position.x; c.position[1] = l.position.y; c.position[2] = l.position.z;
        c.direction[0] = l.direction.x; c.direction[1] = l.direction.y; c.direction[2] = l.direction.z;
        c.color[0] = l.color.x; c.color[1] = l.color.y; c.color[2] = l.color.z;
        c.intensity = l.intensity;
        c.range = l.range;
        c.inner_cone = l.innerCone;
        c.outer_cone = l.outerCone;
        st.lights.push_back(c);
    }
    for (const auto& im : scene.images) {
        std::vector<float> px(4 * im.pixels.size());
        for (size_t i = 0; i < im.pixels.size(); ++i) {
            px[4 * i] = im.pixels[i].x; px[4 * i + 1] = im.pixels[i].y;
            px[4 * i + 2] = im.pixels[i].z; px[4 * i + 3] = im.pixels[i].w;
        }
        st.pixels.push_back(std::move(px));
        st.images.push_back(crg_image{im.width, im.height, st.pixels.back().data()});
    }
    for (const auto& t : scene.textures)
        st.textures.push_back(crg_texture{t.image, static_cast<int32_t>(t.wrapS), static_cast<int32_t>(t.wrapT),
                                          t.linear ? 1 : 0});
    crg_scene_desc d{};
    d.triangle_count = n;
    d.positions = st.positions.data();
    d.normals = st.normals.data();
    d.tangents = st.tangents.data();
    d.uvs = st.uvs.data();
    d.material_ids = st.material_ids.data();
    d.material_count = uint32_t(st.materials.size());
    d.materials = st.materials.data();
    d.light_count = uint32_t(st.lights.size());
    d.lights = st.lights.data();
    d.image_count = uint32_t(st.images.size());
    d.images = st.images.data();
    d.texture_count = uint32_t(st.textures.size());
    d.textures = st.textures.data();
    return d;
}

template <class Camera>
crg_camera camera_of(const Camera& c) {
    crg_camera k{};
    const auto put = [](float* o, auto v) { o[0] = v.x; o[1] = v.y; o[2] = v.z; };
    put(k.position, c.position);
    put(k.forward, c.forward);
    put(k.right, c.right);
    put(k.up, c.up);
    k.yfov = c.yfov;
    k.aspect = c.aspect;
    k.near_clip = c.nearClip;
    k.far_clip = c.farClip;
    return k;
}

template <class Options>
crg_options options_of(const Options& options) {
    crg_options o{};
    o.width = options.width;
    o.height = options.height;
    o.samples = options.samples;
    o.bounces = options.bounces;
    o.threads = options.threads;
    o.next_event_estimation = options.nextEventEstimation;
    o.russian_roulette = options.russianRoulette;
    o.seed = options.seed;
    return o;
}

// A reference Environment as the C descriptor.  The reference keeps an image
// environment's texels private (reference src/render.h:8-28), so only a
// constant environment (no owned storage: bytes() == 0) converts; its
// radiance is the same in every direction (reference src/render.cpp:53).
template <class Environment>
crg_env_desc env_of(const Environment& env) {
    if (env.bytes() != 0)
        throw std::invalid_argument(
            "cr::gpu::render: an image Environment keeps its texels private; pass the crg_env_desc built from the "
            "same image, rotation and intensity (the render overload taking crg_env_desc)");
    crg_env_desc d{};
    const auto c = env.radiance({0.0f, 1.0f, 0.0f});
    d.constant[0] = c.x;
    d.constant[1] = c.y;
    d.constant[2] = c.z;
    d.intensity = 1;
    return d;
}

// Render through the engine: the request is submitted, then polled every
// 50 ms; the caller's Progress is called with (finished camera samples, total,
// seconds) at most every 2 s and once at the end with (total, total, seconds),
// as the reference reports (reference src/render.cpp:419-427, :454).  Scene
// and environment are uploaded for this call and released after it.
template <class Result, class Scene, class Camera, class Options, class Progress>
Result render_with(Device& dev, const Scene& scene, const Camera& camera, const crg_env_desc& env,
                   const Options& options, Progress& progress) {
    const crg_options o = options_of(options);
    Result result;
    // The film is allocated only for dimensions the reference accepts
    // (reference src/render.cpp:97-100); otherwise the library's validation
    // reports the reference's message.
    if (o.width > 0 && o.height > 0 && o.width <= 16384 && o.height <= 16384 &&
        uint64_t(o.width) * uint64_t(o.height) <= 67108864ull)
        result.pixels.resize(size_t(o.width) * size_t(o.height));
    static_assert(sizeof(result.pixels[0]) == 3 * sizeof(float), "film pixels are three floats");
    SceneStorage storage;
    const crg_scene_desc desc = describe(scene, storage);
    uint64_t sid = 0, eid = 0;
    dev.check(crg_scene_upload(dev.handle(), &desc, &sid, nullptr));
    const int erc = crg_env_upload(dev.handle(), &env, &eid, nullptr);
    if (erc != CRG_OK) {
        const std::string e = crg_last_error(dev.handle());
        crg_scene_release(dev.handle(), sid);
        Device::raise(erc, e);
    }
    crg_request req{};
    req.scene = sid;
    req.env = eid;
    req.camera = camera_of(camera);
    req.options = o;
    req.film_rgb = reinterpret_cast<float*>(result.pixels.data());
    uint64_t ticket = 0;
    int rc = req.film_rgb ? crg_submit(dev.handle(), &req, &ticket)
                          : crg_render(dev.handle(), sid, eid, &req.camera, &o, nullptr, nullptr);
    crg_stats stats{};
    std::string error;
    if (rc == CRG_OK) {
        const auto start = std::chrono::steady_clock::now();
        auto last = start;
        while ((rc = crg_wait(dev.handle(), ticket, 0.05, &stats)) == CRG_EBUSY) {
            const auto now = std::chrono::steady_clock::now();
            if (progress && now - last >= std::chrono::seconds(2)) {
                last = now;
                uint64_t done = 0, total = 0;
                if (crg_request_progress(dev.handle(), ticket, &done, &total) == CRG_OK)
                    progress(done, total, std::chrono::duration<double>(now - start).count());
            }
        }
    }
    if (rc != CRG_OK) error = crg_last_error(dev.handle());  // before the releases replace it
    crg_scene_release(dev.handle(), sid);
    crg_env_release(dev.handle(), eid);
    if (rc != CRG_OK) Device::raise(rc, error);
    result.seconds = stats.seconds;
    result.workers = 0;  // the device, not host worker threads
    result.cameraSamples = stats.camera_samples;
    result.boundaryRetries = stats.boundary_retries;
    if (progress) progress(result.cameraSamples, result.cameraSamples, result.seconds);
    return result;
}

// Drop-in for cr::render (reference src/render.h:53) on a given device, with
// the reference's own Environment (constant environments; an image
// environment throws, see env_of).  include/crgpu_render.hpp adds the overload
// with exactly the reference's parameter list on device 0.
template <class Result, class Scene, class Bvh, class Camera, class Environment, class Options, class Progress>
    requires requires(const Environment& e) {
        e.bytes();
        e.radiance({0.0f, 1.0f, 0.0f});
    }
Result render(Device& dev, const Scene& scene, const Bvh&, const Camera& camera, const Environment& env,
              const Options& options, Progress progress) {
    const crg_env_desc d = env_of(env);
    return render_with<Result>(dev, scene, camera, d, options, progress);
}

// The same with the environment as a descriptor: image environments are given
// by the HDR image, rotation and intensity passed to the reference's
// Environment constructor (reference src/main.cpp:212).
template <class Result, class Scene, class Bvh, class Camera, class Options, class Progress>
Result render(Device& dev, const Scene& scene, const Bvh&, const Camera& camera, const crg_env_desc& env,
              const Options& options, Progress progress) {
    return render_with<Result>(dev, scene, camera, env, options, progress);
}

// Drop-in for Bvh::intersect (reference src/scene.h:126-127) over a scene
// made resident once.  intersect() answers a batch of rays in one launch;
// the single-ray form has the reference's exact argument list.
class ResidentScene {
public:
    template <class Scene>
    ResidentScene(Device& dev, const Scene& scene) : dev_(&dev) {
        SceneStorage storage;
        const crg_scene_desc desc = describe(scene, storage);
        dev.check(crg_scene_upload(dev.handle(), &desc, &id_, nullptr));
    }
    ~ResidentScene() { crg_scene_release(dev_->handle(), id_); }
    ResidentScene(const ResidentScene&) = delete;
    ResidentScene& operator=(const ResidentScene&) = delete;
    uint64_t id() const { return id_; }

    template <class Ray, class Hit>
    void intersect(const Ray* rays, Hit* hits, bool* found, size_t count, bool any = false,
                   uint32_t ignore = 0xffffffffu) const {
        std::vector<crg_ray> in(count);
        std::vector<crg_hit> out(count);
        for (size_t i = 0; i < count; ++i) {
            const Ray& r = rays[i];
            in[i] = crg_ray{{r.origin.x, r.origin.y, r.origin.z}, {r.direction.x, r.direction.y, r.direction.z},
                            r.tmin, r.tmax, {r.originLow.x, r.originLow.y, r.originLow.z}, ignore};
        }
        dev_->check(crg_trace_rays(dev_->handle(), id_, in.data(), count, out.data(), any ? 1 : 0));
        for (size_t i = 0; i < count; ++i) {
            found[i] = out[i].hit != 0;
            if (found[i]) {
                hits[i].t = out[i].t;
                hits[i].u = out[i].u;
                hits[i].v = out[i].v;
                hits[i].triangle = out[i].triangle;
            }
        }
    }

    template <class Ray, class Hit>
    bool intersect(const Ray& ray, Hit& hit, bool any = false, uint32_t ignore = 0xffffffffu) const {
        bool found = false;
        intersect(&ray, &hit, &found, 1, any, ignore);
        return found;
    }

private:
    Device* dev_;
    uint64_t id_ = 0;
};

}  // namespace cr::gpu
