// capi.cpp — the crgpu C ABI (include/crgpu.h) over the gfx950 kernels.
//
// A context owns one HIP stream on one device, the resident scenes and
// environments uploaded to it, and a pool of path slots that every render on
// the context reuses.  Nothing here throws across the ABI: failures become a
// status code plus crg_last_error() text in the reference's wording.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <chrono>
#include <cstring>
#include <map>
#include <memory>
#include <mutex>
#include <sstream>
#include <string>
#include <vector>

#include "../../include/crgpu.h"
#include "host_scene.h"
#include "kernels.h"

using namespace cr;

namespace {

struct HipError : std::runtime_error {
    using std::runtime_error::runtime_error;
};
// A transport failure inside a render (reference src/render.cpp:429-451).
struct PathError : std::runtime_error {
    using std::runtime_error::runtime_error;
};

void check(hipError_t e, const char* what) {
    if (e != hipSuccess) {
        if (e == hipErrorOutOfMemory) throw std::bad_alloc();
        throw HipError(std::string(what) + ": " + hipGetErrorString(e));
    }
}

struct DevBuf {
    void* p = nullptr;
    size_t bytes = 0;
    DevBuf() = default;
    DevBuf(const DevBuf&) = delete;
    DevBuf& operator=(const DevBuf&) = delete;
    ~DevBuf() { release(); }
    void release() {
        if (p) (void)hipFree(p);
        p = nullptr;
        bytes = 0;
    }
    void alloc(size_t n) {
        if (n <= bytes && p) return;
        release();
        check(hipMalloc(&p, std::max<size_t>(n, 16)), "hipMalloc");
        bytes = std::max<size_t>(n, 16);
    }
    template <class T>
    T* as() const { return static_cast<T*>(p); }
};

template <class T>
void upload(DevBuf& b, const std::vector<T>& v, hipStream_t st) {
    b.alloc(v.size() * sizeof(T));
    if (!v.empty()) check(hipMemcpyAsync(b.p, v.data(), v.size() * sizeof(T), hipMemcpyHostToDevice, st), "upload");
}

struct SceneRes {
    DevBuf nodes, prims, prim_ng, aux, materials, light_rank, punctual, desc;
    DevScene host{};
    SceneLights lights;
    std::vector<uint32_t> tri_to_prim;
    std::vector<uint32_t> prim_to_tri;
    uint64_t triangles = 0, bytes = 0;
};

struct EnvRes {
    DevBuf texels, pdf, threshold, guide, rowcos, desc;
    DevEnv host{};
    double integral = 0;
    uint64_t bytes = 0;
};

struct Pool {
    uint32_t size = 0;
    DevBuf ray_o, ray_d, ray_lo, thr, rad, lastpos, absorb, rng, job, pixel, sample, medium, flags, prev, hit;
    DevBuf sh_path, sh_o, sh_d, sh_lo, sh_c;
    DevBuf q0, q1, retire, xq, xsq;
    PathSoA P{};
    ShadowSoA S{};
    Queues Q{};

    void ensure(uint32_t n) {
        if (n <= size) return;
        const size_t f4 = size_t(n) * sizeof(float4), u32 = size_t(n) * 4;
        ray_o.alloc(f4); ray_d.alloc(f4); ray_lo.alloc(f4); thr.alloc(f4); rad.alloc(f4); lastpos.alloc(f4);
        absorb.alloc(f4); rng.alloc(size_t(n) * 8); job.alloc(u32); pixel.alloc(u32); sample.alloc(u32);
        medium.alloc(u32); flags.alloc(u32); prev.alloc(u32); hit.alloc(u32);
        sh_path.alloc(u32); sh_o.alloc(f4); sh_d.alloc(f4); sh_lo.alloc(f4); sh_c.alloc(f4);
        q0.alloc(u32); q1.alloc(u32); retire.alloc(u32); xq.alloc(u32); xsq.alloc(u32);
        P = PathSoA{ray_o.as<float4>(), ray_d.as<float4>(), ray_lo.as<float4>(), thr.as<float4>(), rad.as<float4>(),
                    lastpos.as<float4>(), absorb.as<float4>(), rng.as<uint64_t>(), job.as<uint32_t>(),
                    pixel.as<uint32_t>(), sample.as<uint32_t>(), medium.as<int32_t>(), flags.as<uint32_t>(),
                    prev.as<uint32_t>(), hit.as<uint32_t>()};
        S = ShadowSoA{sh_path.as<uint32_t>(), sh_o.as<float4>(), sh_d.as<float4>(), sh_lo.as<float4>(),
                      sh_c.as<float4>()};
        Q = Queues{{q0.as<uint32_t>(), q1.as<uint32_t>()}, retire.as<uint32_t>(), xq.as<uint32_t>(),
                   xsq.as<uint32_t>()};
        size = n;
    }
};

}  // namespace

// Event-pair timing of kernel launches on the context stream.
struct Profiler {
    static constexpr int kSlots = kStageCount + 2;  // + seed, accumulate
    static constexpr int kPairs = 256;
    bool on = false;
    std::vector<hipEvent_t> ev;
    int stage_of[kPairs] = {};
    int used = 0;
    double total_ms[kSlots] = {};
    uint64_t launches[kSlots] = {};

    ~Profiler() {
        for (hipEvent_t e : ev) (void)hipEventDestroy(e);
    }
    void begin(int slot, hipStream_t st) {
        if (!on) return;
        if (ev.empty()) {
            ev.resize(2 * kPairs);
            for (auto& e : ev) check(hipEventCreate(&e), "hipEventCreate");
        }
        if (used == kPairs) drain(st);
        stage_of[used] = slot;
        check(hipEventRecord(ev[2 * used], st), "hipEventRecord");
    }
    void end(hipStream_t st) {
        if (!on) return;
        check(hipEventRecord(ev[2 * used + 1], st), "hipEventRecord");
        ++used;
    }
    void drain(hipStream_t st) {
        if (!used) return;
        check(hipStreamSynchronize(st), "profiler sync");
        for (int i = 0; i < used; ++i) {
            float ms = 0;
            check(hipEventElapsedTime(&ms, ev[2 * i], ev[2 * i + 1]), "hipEventElapsedTime");
            total_ms[stage_of[i]] += ms;
            ++launches[stage_of[i]];
        }
        used = 0;
    }
    static const char* name(int slot) {
        if (slot < kStageCount) return kStageNames[slot];
        return slot == kStageCount ? "k_seed" : "k_accumulate";
    }
};

struct crg_ctx {
    int device = 0;
    Profiler prof;
    hipStream_t stream = nullptr;
    std::string error;
    std::mutex mu;
    uint64_t next_id = 1;
    std::map<uint64_t, std::unique_ptr<SceneRes>> scenes;
    std::map<uint64_t, std::unique_ptr<EnvRes>> envs;
    Pool pool;
    DevBuf ctl, jobs, lights, samples, film, rays, hits;
    uint32_t pool_size = 1u << 22;  // 4M in-flight paths
    unsigned grid = 0;
};

namespace {

template <class F>
int guarded(crg_ctx* ctx, F&& f) {
    if (!ctx) return CRG_EINVAL;
    std::lock_guard<std::mutex> lock(ctx->mu);
    try {
        f();
        ctx->error.clear();
        return CRG_OK;
    } catch (const std::bad_alloc&) {
        ctx->error = "out of memory";
        return CRG_ENOMEM;
    } catch (const HipError& e) {
        ctx->error = e.what();
        return CRG_EDEVICE;
    } catch (const PathError& e) {
        ctx->error = e.what();
        return CRG_EPATH;
    } catch (const std::out_of_range& e) {
        ctx->error = e.what();
        return CRG_ENOTFOUND;
    } catch (const std::exception& e) {
        ctx->error = e.what();
        return CRG_EINVAL;
    }
}

}  // namespace

extern "C" {

const char* crg_version(void) { return "crgpu 0.1 gfx950 wavefront (binary64-exact predicates, no FP64 instructions)"; }

int crg_context_create(int device, crg_ctx** out) {
    if (!out) return CRG_EINVAL;
    *out = nullptr;
    int count = 0;
    if (hipGetDeviceCount(&count) != hipSuccess || count <= device || device < 0) return CRG_EDEVICE;
    hipDeviceProp_t prop;
    if (hipGetDeviceProperties(&prop, device) != hipSuccess) return CRG_EDEVICE;
    if (std::string(prop.gcnArchName).find("gfx950") == std::string::npos) return CRG_EDEVICE;
    auto* ctx = new (std::nothrow) crg_ctx();
    if (!ctx) return CRG_ENOMEM;
    ctx->device = device;
    if (hipSetDevice(device) != hipSuccess || hipStreamCreateWithFlags(&ctx->stream, hipStreamNonBlocking) != hipSuccess) {
        delete ctx;
        return CRG_EDEVICE;
    }
    ctx->grid = unsigned(prop.multiProcessorCount) * 8;
    *out = ctx;
    return CRG_OK;
}

void crg_context_destroy(crg_ctx* ctx) {
    if (!ctx) return;
    (void)hipSetDevice(ctx->device);
    (void)hipStreamSynchronize(ctx->stream);
    ctx->scenes.clear();
    ctx->envs.clear();
    (void)hipStreamDestroy(ctx->stream);
    delete ctx;
}

const char* crg_last_error(const crg_ctx* ctx) { return ctx ? ctx->error.c_str() : "null context"; }

int crg_scene_upload(crg_ctx* ctx, const crg_scene_desc* desc, uint64_t* scene_id, crg_scene_info* info) {
    return guarded(ctx, [&] {
        if (!desc || !scene_id) throw std::invalid_argument("null scene descriptor");
        if (desc->triangle_count && (!desc->positions || !desc->material_ids))
            throw std::invalid_argument("scene descriptor without positions or material ids");
        check(hipSetDevice(ctx->device), "hipSetDevice");
        const auto t0 = std::chrono::steady_clock::now();
        HostScene hs = finalize_scene(*desc);
        for (size_t i = 0; i < hs.materials.size(); ++i) {
            const crg_material& m = hs.materials[i];
            for (const crg_texture_slot* s :
                 {&m.base_color_tex, &m.metallic_roughness_tex, &m.normal_tex, &m.emissive_tex, &m.transmission_tex})
                if (s->texture >= 0)
                    throw std::invalid_argument("Material " + std::to_string(i) +
                                                " uses a texture map; the device path supports factor-only materials");
        }
        WideBvh bvh = build_wide_bvh(hs);
        auto res = std::make_unique<SceneRes>();
        const size_t n = hs.tris.size();
        res->triangles = n;
        res->prim_to_tri = bvh.order;
        res->tri_to_prim.assign(n, 0);
        for (size_t i = 0; i < n; ++i) res->tri_to_prim[bvh.order[i]] = uint32_t(i);
        std::vector<PrimRec> prims(n);
        std::vector<float4> png(n);
        std::vector<PrimAux> aux(n);
        for (size_t i = 0; i < n; ++i) {
            const HTriangle& t = hs.tris[bvh.order[i]];
            const crg_material& m = hs.materials[t.material];
            uint32_t flags = 0;
            if (!m.double_sided && m.transmission <= 0) flags |= kPrimCull;
            if (m.alpha_mode == 1 && !(m.base_color[3] * 1.0f >= m.alpha_cutoff)) flags |= kPrimNeverHit;
            prims[i] = PrimRec{t.p[0].x, t.p[0].y, t.p[0].z, bvh.order[i], t.p[1].x, t.p[1].y, t.p[1].z, flags,
                               t.p[2].x, t.p[2].y, t.p[2].z, t.material};
            png[i] = make_float4(t.ng.x, t.ng.y, t.ng.z, 0);
            PrimAux& a = aux[i];
            const HVec3* nv[3] = {&t.n[0], &t.n[1], &t.n[2]};
            float* dst[3] = {a.n0, a.n1, a.n2};
            for (int k = 0; k < 3; ++k) {
                dst[k][0] = nv[k]->x;
                dst[k][1] = nv[k]->y;
                dst[k][2] = nv[k]->z;
            }
            a.ng[0] = t.ng.x; a.ng[1] = t.ng.y; a.ng[2] = t.ng.z;
            a.ung[0] = t.ung.x; a.ung[1] = t.ung.y; a.ung[2] = t.ung.z;
            a.area = t.area;
        }
        res->lights = scene_lights(hs);
        const std::vector<uint32_t> ranks = light_ranks(res->lights, res->tri_to_prim, n);
        const std::vector<DevMaterial> mats = device_materials(hs);
        const auto t1 = std::chrono::steady_clock::now();
        hipStream_t st = ctx->stream;
        upload(res->nodes, bvh.nodes, st);
        upload(res->prims, prims, st);
        upload(res->prim_ng, png, st);
        upload(res->aux, aux, st);
        upload(res->materials, mats, st);
        upload(res->light_rank, ranks, st);
        upload(res->punctual, res->lights.punctual, st);
        res->host = DevScene{res->nodes.as<WNode>(), res->prims.as<PrimRec>(), res->prim_ng.as<float4>(),
                             res->aux.as<PrimAux>(), res->materials.as<DevMaterial>(), res->light_rank.as<uint32_t>(),
                             res->punctual.as<DevPunctual>(), uint32_t(bvh.nodes.size()), uint32_t(n),
                             uint32_t(res->lights.punctual.size()), 0};
        res->desc.alloc(sizeof(DevScene));
        check(hipMemcpyAsync(res->desc.p, &res->host, sizeof(DevScene), hipMemcpyHostToDevice, st), "upload");
        check(hipStreamSynchronize(st), "scene upload");
        const auto t2 = std::chrono::steady_clock::now();
        res->bytes = res->nodes.bytes + res->prims.bytes + res->prim_ng.bytes + res->aux.bytes + res->materials.bytes +
                     res->light_rank.bytes + res->punctual.bytes + res->desc.bytes;
        if (info) {
            info->triangles = n;
            info->bvh_nodes = bvh.binary_nodes;
            info->wide_nodes = bvh.nodes.size();
            info->device_bytes = res->bytes;
            info->build_seconds = std::chrono::duration<double>(t1 - t0).count();
            info->upload_seconds = std::chrono::duration<double>(t2 - t1).count();
        }
        const uint64_t id = ctx->next_id++;
        ctx->scenes[id] = std::move(res);
        *scene_id = id;
    });
}

int crg_scene_release(crg_ctx* ctx, uint64_t scene_id) {
    return guarded(ctx, [&] {
        (void)hipStreamSynchronize(ctx->stream);
        if (!ctx->scenes.erase(scene_id)) throw std::out_of_range("unknown scene handle");
    });
}

int crg_env_upload(crg_ctx* ctx, const crg_env_desc* desc, uint64_t* env_id, uint64_t* device_bytes) {
    return guarded(ctx, [&] {
        if (!desc || !env_id) throw std::invalid_argument("null environment descriptor");
        check(hipSetDevice(ctx->device), "hipSetDevice");
        HostEnv he = build_env(*desc);
        auto res = std::make_unique<EnvRes>();
        hipStream_t st = ctx->stream;
        res->integral = he.integral;
        DevEnv& d = res->host;
        d = DevEnv{};
        d.width = he.width;
        d.height = he.height;
        d.rotation = he.rotation;
        d.intensity = he.intensity;
        std::memcpy(d.constant, he.constant, sizeof(d.constant));
        d.is_image = he.is_image ? 1 : 0;
        if (he.is_image) {
            upload(res->texels, he.texels, st);
            upload(res->pdf, he.pdf, st);
            upload(res->threshold, he.threshold, st);
            upload(res->guide, he.guide, st);
            upload(res->rowcos, he.rowcos, st);
            d.texels = res->texels.as<float4>();
            d.pdf = res->pdf.as<float>();
            d.threshold = res->threshold.as<uint32_t>();
            d.guide = res->guide.as<uint32_t>();
            d.rowcos = res->rowcos.as<float>();
        }
        res->desc.alloc(sizeof(DevEnv));
        check(hipMemcpyAsync(res->desc.p, &d, sizeof(DevEnv), hipMemcpyHostToDevice, st), "upload");
        check(hipStreamSynchronize(st), "environment upload");
        res->bytes = res->texels.bytes + res->pdf.bytes + res->threshold.bytes + res->guide.bytes + res->rowcos.bytes +
                     res->desc.bytes;
        if (device_bytes) *device_bytes = res->bytes;
        const uint64_t id = ctx->next_id++;
        ctx->envs[id] = std::move(res);
        *env_id = id;
    });
}

int crg_env_release(crg_ctx* ctx, uint64_t env_id) {
    return guarded(ctx, [&] {
        (void)hipStreamSynchronize(ctx->stream);
        if (!ctx->envs.erase(env_id)) throw std::out_of_range("unknown environment handle");
    });
}

int crg_trace_rays(crg_ctx* ctx, uint64_t scene_id, const crg_ray* rays, uint64_t count, crg_hit* hits, int any) {
    return guarded(ctx, [&] {
        auto it = ctx->scenes.find(scene_id);
        if (it == ctx->scenes.end()) throw std::out_of_range("unknown scene handle");
        if (count && (!rays || !hits)) throw std::invalid_argument("null ray or hit array");
        check(hipSetDevice(ctx->device), "hipSetDevice");
        hipStream_t st = ctx->stream;
        ctx->rays.alloc(count * sizeof(crg_ray));
        ctx->hits.alloc(count * sizeof(crg_hit));
        ctx->ctl.alloc(sizeof(Control));
        check(hipMemsetAsync(ctx->ctl.p, 0, sizeof(Control), st), "memset");
        if (count) check(hipMemcpyAsync(ctx->rays.p, rays, count * sizeof(crg_ray), hipMemcpyHostToDevice, st), "copy");
        check(launch_trace_rays(it->second->desc.as<DevScene>(), ctx->rays.as<RayIn>(), count, ctx->hits.as<HitOut>(),
                                any, ctx->ctl.as<Control>(), st),
              "trace launch");
        if (count) check(hipMemcpyAsync(hits, ctx->hits.p, count * sizeof(crg_hit), hipMemcpyDeviceToHost, st), "copy");
        check(hipStreamSynchronize(st), "trace");
        for (uint64_t i = 0; i < count; ++i)
            if (hits[i].hit == 2u) throw std::logic_error("BVH traversal exceeded its bounded stack");
    });
}

int crg_render(crg_ctx* ctx, uint64_t scene_id, uint64_t env_id, const crg_camera* camera, const crg_options* options,
               float* film_rgb, crg_stats* stats) {
    return guarded(ctx, [&] {
        if (!camera || !options || !film_rgb) throw std::invalid_argument("null render argument");
        const crg_options o = *options;
        validate_options(o);
        auto si = ctx->scenes.find(scene_id);
        if (si == ctx->scenes.end()) throw std::out_of_range("unknown scene handle");
        auto ei = ctx->envs.find(env_id);
        if (ei == ctx->envs.end()) throw std::out_of_range("unknown environment handle");
        SceneRes& sc = *si->second;
        EnvRes& env = *ei->second;
        check(hipSetDevice(ctx->device), "hipSetDevice");
        hipStream_t st = ctx->stream;
        DevJob job{};
        job.cam = prepare_camera(*camera, o.width, o.height);
        const LightTable lt = build_light_table(sc.lights, sc.tri_to_prim, env.integral);
        upload(ctx->lights, lt.entries, st);
        job.scene = sc.desc.as<DevScene>();
        job.env = env.desc.as<DevEnv>();
        job.lights = DevLights{ctx->lights.as<DevLightEntry>(), uint32_t(lt.entries.size()), lt.env_probability};
        job.width = o.width;
        job.height = o.height;
        job.spp = o.samples;
        job.bounces = o.bounces;
        job.seed = o.seed;
        job.nee = o.next_event_estimation;
        job.rr = o.russian_roulette;
        const uint64_t npix = uint64_t(o.width) * uint64_t(o.height);
        job.items = npix * uint64_t(o.samples);
        job.item_base = 0;
        ctx->samples.alloc(job.items * sizeof(float4));
        job.samples = ctx->samples.as<float4>();
        ctx->jobs.alloc(sizeof(DevJob));
        check(hipMemcpyAsync(ctx->jobs.p, &job, sizeof(DevJob), hipMemcpyHostToDevice, st), "job upload");
        const uint32_t pool = uint32_t(std::min<uint64_t>(ctx->pool_size, job.items));
        ctx->pool.ensure(pool);
        ctx->ctl.alloc(sizeof(Control));
        ctx->film.alloc(npix * 3 * sizeof(float));
        Control init{};
        init.total = job.items;
        init.issued = pool;
        check(hipMemcpyAsync(ctx->ctl.p, &init, sizeof(Control), hipMemcpyHostToDevice, st), "control upload");
        check(hipStreamSynchronize(st), "render setup");
        Control* dctl = ctx->ctl.as<Control>();
        const DevJob* djobs = ctx->jobs.as<DevJob>();
        const auto start = std::chrono::steady_clock::now();
        Profiler& prof = ctx->prof;
        prof.begin(kStageCount, st);
        check(launch_seed(ctx->pool.P, djobs, 1, pool, ctx->pool.Q.q[0], dctl, st), "seed launch");
        prof.end(st);
        int parity = 0, iterations = 0;
        Control head{};
        const int batch = 4;
        for (;;) {
            for (int k = 0; k < batch; ++k) {
                for (int stage = 0; stage < kStageCount; ++stage) {
                    prof.begin(stage, st);
                    check(launch_stage(Stage(stage), ctx->pool.P, ctx->pool.S, djobs, 1, ctx->pool.Q, parity, dctl,
                                       ctx->grid, st),
                          "iteration launch");
                    prof.end(st);
                }
                parity ^= 1;
                ++iterations;
            }
            check(hipMemcpyAsync(&head, dctl, offsetof(Control, errors), hipMemcpyDeviceToHost, st), "control read");
            check(hipStreamSynchronize(st), "render iteration");
            prof.drain(st);
            if (head.issued >= head.total && head.qcount[parity] == 0) break;
            // Every path finishes within 2*(bounces+1) iterations of its start (one
            // precision replay at most), so this bound only trips on a logic error.
            const uint64_t bound = (job.items / pool + 2) * uint64_t(2 * (o.bounces + 2)) + 64;
            if (uint64_t(iterations) > bound) throw std::logic_error("render did not converge");
        }
        prof.begin(kStageCount + 1, st);
        check(launch_accumulate(djobs, 0, npix, ctx->film.as<float>(), st), "accumulate launch");
        prof.end(st);
        check(hipMemcpyAsync(film_rgb, ctx->film.p, npix * 3 * sizeof(float), hipMemcpyDeviceToHost, st), "film read");
        check(hipStreamSynchronize(st), "render");
        prof.drain(st);
        const double seconds = std::chrono::duration<double>(std::chrono::steady_clock::now() - start).count();
        check(hipMemcpy(&head, dctl, sizeof(Control), hipMemcpyDeviceToHost), "control read");
        if (head.err_count) {
            const uint32_t n = std::min<uint32_t>(head.err_count, kMaxErrors);
            const ErrorRec* best = &head.errors[0];
            for (uint32_t i = 1; i < n; ++i)
                if (head.errors[i].key < best->key) best = &head.errors[i];
            throw PathError(format_path_error(*best, o.width, o.samples));
        }
        if (stats) {
            std::memset(stats, 0, sizeof(*stats));
            stats->seconds = seconds;
            stats->camera_samples = job.items;
            stats->boundary_retries = head.replays;
            stats->extend_rays = head.extend_rays;
            stats->shadow_rays = head.shadow_rays;
            stats->workers = 0;
            stats->iterations = iterations;
            stats->extend_nodes = head.extend_nodes;
            stats->extend_tris = head.extend_tris;
            stats->shadow_nodes = head.shadow_nodes;
            stats->shadow_tris = head.shadow_tris;
            stats->exact_fallbacks = head.fallbacks;
            stats->exact_rays = head.exact_rays;
        }
    });
}

int crg_set_profiling(crg_ctx* ctx, int enable) {
    return guarded(ctx, [&] {
        ctx->prof.on = enable != 0;
        if (enable) {
            std::fill(std::begin(ctx->prof.total_ms), std::end(ctx->prof.total_ms), 0.0);
            std::fill(std::begin(ctx->prof.launches), std::end(ctx->prof.launches), 0ull);
        }
    });
}

int crg_kernel_times(crg_ctx* ctx, crg_kernel_time* out, int max, int* count) {
    return guarded(ctx, [&] {
        const int n = Profiler::kSlots;
        if (count) *count = n;
        for (int i = 0; i < n && i < max && out; ++i) {
            std::memset(&out[i], 0, sizeof(out[i]));
            std::strncpy(out[i].name, Profiler::name(i), sizeof(out[i].name) - 1);
            out[i].launches = ctx->prof.launches[i];
            out[i].total_ms = ctx->prof.total_ms[i];
        }
    });
}

}  // extern "C"
