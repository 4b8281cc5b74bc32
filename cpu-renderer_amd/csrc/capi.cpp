// capi.cpp — the crgpu C ABI (include/crgpu.h) over the gfx950 kernels.
//
// A context owns, on one device, the resident scenes and environments, an
// upload / ray-query stream, and the wavefront engine (engine.h) that every
// render request — plain, batched or served — runs on.  Nothing here throws
// across the ABI: failures become a status code plus crg_last_error() text in
// the reference's wording.
#include <hip/hip_runtime.h>

#include <fcntl.h>
#include <sys/stat.h>
#include <unistd.h>

#include <algorithm>
#include <atomic>
#include <cerrno>
#include <cstddef>
#include <chrono>
#include <cmath>
#include <cstdio>
#include <cstring>
#include <limits>
#include <map>
#include <memory>
#include <mutex>
#include <sstream>
#include <string>
#include <vector>

#include "../../include/crgpu.h"
#include "device_buf.h"
#include "engine.h"
#include "film.h"
#include "film_host.h"
#include "host_scene.h"
#include "kernels.h"
#include "parallel.h"

using namespace cr;

struct crg_ctx {
    int device = 0;
    hipStream_t stream = nullptr;  // uploads and ray queries
    std::string error;
    std::mutex mu;                 // handles, error text, the upload stream
    uint64_t next_id = 1;
    std::map<uint64_t, std::shared_ptr<SceneRes>> scenes;
    std::map<uint64_t, std::shared_ptr<EnvRes>> envs;
    std::unique_ptr<Engine> engine;
    DevBuf ctl, rays, hits;
    void* scratch = nullptr;       // this device's fixed scratch (crg_trace_rays_noalloc)
    // film encode
    DevBuf film, ldr, planes, streams, lens, thresholds, first_bad;
    float cached_exposure = std::numeric_limits<float>::quiet_NaN();
    float cached_thresholds[255] = {};
};

namespace {

// Runs f() and maps exceptions to status codes; the failure text is kept per
// context.  f() takes the context lock itself where it touches shared state.
// A request's own failure status, carried out of guarded() with its text.
struct StatusError : std::runtime_error {
    int code;
    StatusError(int c, const std::string& m) : std::runtime_error(m), code(c) {}
};

template <class F>
int run_mapped(F&& f, std::string& msg) {
    int rc = CRG_OK;
    try {
        f();
    } catch (const StatusError& e) {
        rc = e.code;
        msg = e.what();
    } catch (const std::bad_alloc&) {
        rc = CRG_ENOMEM;
        msg = "out of memory";
    } catch (const HipError& e) {
        rc = CRG_EDEVICE;
        msg = e.what();
    } catch (const PathError& e) {
        rc = CRG_EPATH;
        msg = e.what();
    } catch (const std::out_of_range& e) {
        rc = CRG_ENOTFOUND;
        msg = e.what();
    } catch (const std::exception& e) {
        rc = CRG_EINVAL;
        msg = e.what();
    }
    return rc;
}

template <class F>
int guarded(crg_ctx* ctx, F&& f) {
    if (!ctx) return CRG_EINVAL;
    std::string msg;
    const int rc = run_mapped(std::forward<F>(f), msg);
    std::lock_guard<std::mutex> lock(ctx->mu);
    ctx->error = msg;
    return rc;
}

// Entry points without a context report their failure text into a caller
// buffer (truncated to capacity, always terminated when capacity > 0).
template <class F>
int guarded_msg(char* message, size_t capacity, F&& f) {
    std::string msg;
    const int rc = run_mapped(std::forward<F>(f), msg);
    if (message && capacity) {
        const size_t n = std::min(capacity - 1, msg.size());
        std::memcpy(message, msg.data(), n);
        message[n] = 0;
    }
    return rc;
}

// Request from C arguments: option/camera validation and the light table,
// on the calling thread (reference src/render.cpp:356-372).
Request make_request(crg_ctx* ctx, const crg_request& r) {
    validate_options(r.options);  // the reference's messages first, as render() validates before allocating
    if (!r.film_rgb) throw std::invalid_argument("null film buffer");
    Request q;
    {
        std::lock_guard<std::mutex> lock(ctx->mu);
        auto si = ctx->scenes.find(r.scene);
        if (si == ctx->scenes.end()) throw std::out_of_range("unknown scene handle");
        auto ei = ctx->envs.find(r.env);
        if (ei == ctx->envs.end()) throw std::out_of_range("unknown environment handle");
        q.scene = si->second;
        q.env = ei->second;
    }
    q.options = r.options;
    q.camera = prepare_camera(r.camera, r.options.width, r.options.height);
    q.lights = build_light_table(q.scene->lights, q.scene->tri_to_prim, q.env->integral);
    q.film = r.film_rgb;
    q.win[0] = 0;
    q.win[1] = 0;
    q.win[2] = r.options.width;
    q.win[3] = r.options.height;
    return q;
}

// CRG_OK or CRG_EBUSY; a failed request throws StatusError.
int status_of_wait(Engine& eng, uint64_t ticket, double timeout, crg_stats* stats) {
    std::string err;
    crg_stats st{};
    const int rc = eng.wait(ticket, timeout, &st, &err);
    if (stats && rc != CRG_EBUSY) *stats = st;
    if (rc != CRG_OK && rc != CRG_EBUSY) throw StatusError(rc, err);
    return rc;
}

// One array of device-format records in host memory (a vector, or a section
// of a scene cache file, which need not be aligned for the record type).
struct Section {
    const void* p = nullptr;
    size_t count = 0;
    size_t bytes = 0;
};
template <class T>
Section section(const std::vector<T>& v) {
    return {v.data(), v.size(), v.size() * sizeof(T)};
}

// The device-format arrays of a scene, wherever they are held on the host.
struct SceneSections {
    uint64_t binary_nodes = 0;
    Section nodes4, nodes8, prims, prim_ng, aux, materials, uvs, tangents, texels, images, textures;
};

SceneSections sections_of(const SceneArrays& a) {
    return {a.bvh.binary_nodes, section(a.bvh.q4), section(a.bvh.q8), section(a.prims),  section(a.prim_ng),
            section(a.aux),     section(a.materials), section(a.uvs),    section(a.tangents),
            section(a.texels),  section(a.images),    section(a.textures)};
}

void upload_section(DevBuf& b, const Section& s, hipStream_t st) {
    b.alloc(s.bytes);
    if (s.bytes) check(hipMemcpyAsync(b.p, s.p, s.bytes, hipMemcpyHostToDevice, st), "upload");
}

// Uploads the device-format arrays of a scene; the caller holds ctx->mu.
// prim_to_tri / tri_to_prim map between leaf order and input order.
std::shared_ptr<SceneRes> make_resident(crg_ctx* ctx, const SceneSections& a, std::vector<uint32_t> prim_to_tri,
                                        std::vector<uint32_t> tri_to_prim, SceneLights lights, uint64_t n) {
    auto res = std::make_shared<SceneRes>();
    res->triangles = n;
    res->binary_nodes = a.binary_nodes;
    res->prim_to_tri = std::move(prim_to_tri);
    res->tri_to_prim = std::move(tri_to_prim);
    res->lights = std::move(lights);
    res->texel_count = a.texels.count;
    res->image_count = a.images.count;
    res->texture_count = a.textures.count;
    res->material_count = a.materials.count;
    const std::vector<uint32_t> ranks = light_ranks(res->lights, res->prim_to_tri, n);
    // Alpha-mask primitives select the traversal variant with FP32 mask tests.
    std::atomic<bool> masks{false};
    parallel_chunks(a.prims.count, size_t(1) << 18, [&](size_t b, size_t e, size_t) {
        const uint8_t* base = static_cast<const uint8_t*>(a.prims.p) + offsetof(PrimRec, flags);
        for (size_t i = b; i < e; ++i) {
            uint32_t f;
            std::memcpy(&f, base + i * sizeof(PrimRec), sizeof(f));
            if (f & kPrimMaskTexture) {
                masks = true;
                return;
            }
        }
    });
    res->masks = masks;
    hipStream_t st = ctx->stream;
    upload_section(res->nodes4, a.nodes4, st);
    upload_section(res->nodes8, a.nodes8, st);
    upload_section(res->prims, a.prims, st);
    upload_section(res->prim_ng, a.prim_ng, st);
    upload_section(res->aux, a.aux, st);
    upload_section(res->materials, a.materials, st);
    upload(res->light_rank, ranks, st);
    upload(res->punctual, res->lights.punctual, st);
    upload_section(res->uvs, a.uvs, st);
    upload_section(res->tangents, a.tangents, st);
    upload_section(res->texels, a.texels, st);
    upload_section(res->images, a.images, st);
    upload_section(res->textures, a.textures, st);
    res->host = DevScene{res->nodes4.as<QNode4>(),
                         res->nodes8.as<QNode8>(),
                         res->prims.as<PrimRec>(),
                         res->prim_ng.as<float4>(),
                         res->aux.as<PrimAux>(),
                         res->materials.as<DevMaterial>(),
                         res->light_rank.as<uint32_t>(),
                         res->punctual.as<DevPunctual>(),
                         a.uvs.count ? res->uvs.as<float4>() : nullptr,
                         a.tangents.count ? res->tangents.as<float4>() : nullptr,
                         res->texels.as<float4>(),
                         res->images.as<DevImage>(),
                         res->textures.as<DevTexture>(),
                         uint32_t(a.nodes4.count),
                         uint32_t(a.prims.count),
                         uint32_t(res->lights.punctual.size()),
                         uint32_t(a.nodes8.count)};
    res->desc.alloc(sizeof(DevScene));
    check(hipMemcpyAsync(res->desc.p, &res->host, sizeof(DevScene), hipMemcpyHostToDevice, st), "upload");
    check(hipStreamSynchronize(st), "scene upload");
    res->bytes = res->nodes4.bytes + res->nodes8.bytes + res->prims.bytes + res->prim_ng.bytes + res->aux.bytes + res->materials.bytes +
                 res->light_rank.bytes + res->punctual.bytes + res->uvs.bytes + res->tangents.bytes +
                 res->texels.bytes + res->images.bytes + res->textures.bytes + res->desc.bytes;
    return res;
}

crg_scene_info scene_info(const SceneRes& r) {
    crg_scene_info info{};
    info.triangles = r.triangles;
    info.bvh_nodes = r.binary_nodes;
    info.wide_nodes = r.host.node_count4;
    info.device_bytes = r.bytes;
    return info;
}

// ---- scene cache file: the device-format arrays of a resident scene, so a
// server can make a scene resident again without finalize and BVH build.
// The magic's last two characters are the format version: version 2 files are
// those whose every index the loader checks before anything is uploaded.
constexpr char kCacheMagic[8] = {'C', 'R', 'G', 'S', 'C', 'N', '0', '2'};

struct CacheWriter {
    std::vector<uint8_t> out;
    void raw(const void* p, size_t n) {
        const uint8_t* b = static_cast<const uint8_t*>(p);
        out.insert(out.end(), b, b + n);
    }
    template <class T>
    void pod(const T& v) { raw(&v, sizeof(T)); }
    template <class T>
    void vec(const std::vector<T>& v) {
        pod(uint64_t(v.size()));
        raw(v.data(), v.size() * sizeof(T));
    }
};

// Reads the sections of a cache file held in memory; `section` returns a view
// into the file's bytes, `vec` a copy.
struct CacheReader {
    const uint8_t* in;
    size_t size;
    size_t pos;
    std::string label;
    void raw(void* p, size_t n) {
        if (n > size - pos) throw std::runtime_error(label + ": truncated");
        std::memcpy(p, in + pos, n);
        pos += n;
    }
    template <class T>
    T pod() {
        T v;
        raw(&v, sizeof(T));
        return v;
    }
    template <class T>
    Section section() {
        const uint64_t n = pod<uint64_t>();
        if (n > (size - pos) / sizeof(T)) throw std::runtime_error(label + ": truncated");
        const Section s{in + pos, size_t(n), size_t(n) * sizeof(T)};
        pos += s.bytes;
        return s;
    }
    template <class T>
    std::vector<T> vec() {
        const Section s = section<T>();
        std::vector<T> v(s.count);
        if (s.bytes) std::memcpy(v.data(), s.p, s.bytes);
        return v;
    }
};

// Checksum of a cache file's body: 1 MiB blocks hashed in parallel, each as
// four interleaved lanes of an FNV-1a-style step over 64-bit words (xor,
// multiply by the FNV prime, rotate: a bijection of the lane for each word, so
// any change to one word changes the block's hash), tail bytes in lane 0; then
// FNV-1a over the byte count and the block hashes.
constexpr uint64_t kFnvBasis = 0xcbf29ce484222325ull, kFnvPrime = 0x100000001b3ull;
constexpr size_t kHashBlock = size_t(1) << 20;

uint64_t fnv_step(uint64_t h, uint64_t w) {
    h = (h ^ w) * kFnvPrime;
    return (h << 23) | (h >> 41);
}

uint64_t block_hash(const uint8_t* p, size_t n) {
    uint64_t h[4] = {kFnvBasis, kFnvBasis + 1, kFnvBasis + 2, kFnvBasis + 3};
    size_t i = 0;
    for (; i + 32 <= n; i += 32) {
        uint64_t w[4];
        std::memcpy(w, p + i, 32);
        for (int k = 0; k < 4; ++k) h[k] = fnv_step(h[k], w[k]);
    }
    for (; i < n; ++i) h[0] = fnv_step(h[0], p[i]);
    uint64_t r = kFnvBasis;
    for (int k = 0; k < 4; ++k) r = fnv_step(r, h[k]);
    return r;
}

uint64_t cache_checksum(const uint8_t* p, size_t n) {
    const size_t blocks = (n + kHashBlock - 1) / kHashBlock;
    std::vector<uint64_t> bh(blocks);
    parallel_chunks(blocks, 8, [&](size_t b, size_t e, size_t) {
        for (size_t i = b; i < e; ++i) bh[i] = block_hash(p + i * kHashBlock, std::min(kHashBlock, n - i * kHashBlock));
    });
    uint64_t r = fnv_step(kFnvBasis, uint64_t(n));
    for (uint64_t h : bh) r = fnv_step(r, h);
    return r;
}

// A whole file in host memory, read by parallel positioned reads.
struct FileBytes {
    std::unique_ptr<uint8_t[]> p;
    size_t n = 0;
};

FileBytes read_file(const char* path, const std::string& label) {
    const int fd = ::open(path, O_RDONLY | O_CLOEXEC);
    if (fd < 0) throw std::runtime_error(label + ": cannot open");
    struct stat st;
    if (::fstat(fd, &st) != 0 || !S_ISREG(st.st_mode)) {
        ::close(fd);
        throw std::runtime_error(label + ": cannot open");
    }
    FileBytes f;
    f.n = size_t(st.st_size);
    f.p.reset(new uint8_t[std::max<size_t>(f.n, 1)]);
    std::atomic<bool> ok{true};
    parallel_chunks(f.n, size_t(16) << 20, [&](size_t b, size_t e, size_t) {
        while (b < e) {
            const ssize_t k = ::pread(fd, f.p.get() + b, std::min<size_t>(e - b, size_t(1) << 30), off_t(b));
            if (k < 0 && errno == EINTR) continue;
            if (k <= 0) {
                ok = false;
                return;
            }
            b += size_t(k);
        }
    });
    ::close(fd);
    if (!ok) throw std::runtime_error(label + ": truncated");
    return f;
}

const uint32_t kRecordSizes[9] = {sizeof(QNode4), sizeof(QNode8),     sizeof(PrimRec),
                                  sizeof(PrimAux), sizeof(DevMaterial), sizeof(DevPunctual),
                                  sizeof(DevImage), sizeof(DevTexture), sizeof(float4)};

template <class T>
std::vector<T> download(const DevBuf& b, uint64_t count, hipStream_t st) {
    std::vector<T> v(count);
    if (count) check(hipMemcpyAsync(v.data(), b.p, count * sizeof(T), hipMemcpyDeviceToHost, st), "download");
    return v;
}

// The device-format arrays of a scene downloaded from the device.
struct CacheContent {
    std::vector<QNode4> nodes4;
    std::vector<QNode8> nodes8;
    std::vector<PrimRec> prims;
    std::vector<float4> prim_ng;
    std::vector<PrimAux> aux;
    std::vector<DevMaterial> materials;
    std::vector<float4> uvs, tangents, texels;
    std::vector<DevImage> images;
    std::vector<DevTexture> textures;
};

SceneSections sections_of(const CacheContent& c) {
    return {0,
            section(c.nodes4),   section(c.nodes8), section(c.prims),  section(c.prim_ng), section(c.aux),
            section(c.materials),
            section(c.uvs),      section(c.tangents), section(c.texels), section(c.images), section(c.textures)};
}

std::string cache_label(const char* path) { return std::string("Scene cache '") + path + "'"; }

void put_section(CacheWriter& w, const Section& s) {
    w.pod(uint64_t(s.count));
    w.raw(s.p, s.bytes);
}

// Writes a cache file: magic, record sizes, triangle and binary-node counts,
// the device-format sections, the leaf-order maps, the light tables, and the
// checksum of everything before it.
void write_cache_file(const char* path, uint64_t triangles, uint64_t binary_nodes, const SceneSections& sec,
                      const std::vector<uint32_t>& prim_to_tri, const std::vector<uint32_t>& tri_to_prim,
                      const SceneLights& lights) {
    CacheWriter w;
    w.out.reserve(4096 + sec.nodes4.bytes + sec.nodes8.bytes + sec.prims.bytes + sec.prim_ng.bytes + sec.aux.bytes + sec.materials.bytes +
                  sec.uvs.bytes + sec.tangents.bytes + sec.texels.bytes + sec.images.bytes + sec.textures.bytes +
                  8 * (prim_to_tri.size() + tri_to_prim.size()));
    w.raw(kCacheMagic, 8);
    w.raw(kRecordSizes, sizeof(kRecordSizes));
    w.pod(triangles);
    w.pod(binary_nodes);
    for (const Section* s : {&sec.nodes4, &sec.nodes8, &sec.prims, &sec.prim_ng, &sec.aux, &sec.materials, &sec.uvs, &sec.tangents,
                             &sec.texels, &sec.images, &sec.textures})
        put_section(w, *s);
    w.vec(prim_to_tri);
    w.vec(tri_to_prim);
    w.pod(lights.radius2);
    w.vec(lights.punctual_weight);
    w.vec(lights.emissive);
    w.vec(lights.emissive_weight);
    w.vec(lights.punctual);
    w.pod(cache_checksum(w.out.data(), w.out.size()));
    std::FILE* f = std::fopen(path, "wb");
    const bool ok = f && std::fwrite(w.out.data(), 1, w.out.size(), f) == w.out.size();
    if (!f || std::fclose(f) != 0 || !ok) throw std::runtime_error(std::string("Failed writing scene cache '") + path + "'");
}

// A cache file's sections as views into its bytes (nodes ... textures) and
// copies (maps, lights).
struct ParsedCache {
    uint64_t triangles = 0;
    SceneSections sec;
    std::vector<uint32_t> order, tri_to_prim;
    SceneLights lights;
};

ParsedCache parse_cache(const FileBytes& file, const std::string& label) {
    const uint8_t* bytes = file.p.get();
    if (file.n < 8 + sizeof(kRecordSizes) + 24 || std::memcmp(bytes, kCacheMagic, 6) != 0)
        throw std::runtime_error(label + ": not a scene cache of this library");
    if (std::memcmp(bytes, kCacheMagic, 8) != 0)
        throw std::runtime_error(label + ": written in another format version of this library");
    uint64_t stored;
    std::memcpy(&stored, bytes + file.n - 8, 8);
    if (stored != cache_checksum(bytes, file.n - 8)) throw std::runtime_error(label + ": checksum mismatch");
    CacheReader rd{bytes, file.n - 8, 8, label};
    uint32_t sizes[9];
    rd.raw(sizes, sizeof(sizes));
    if (std::memcmp(sizes, kRecordSizes, sizeof(sizes)) != 0)
        throw std::runtime_error(label + ": written by a build with other record layouts");
    ParsedCache pc;
    pc.triangles = rd.pod<uint64_t>();
    SceneSections& sec = pc.sec;
    sec.binary_nodes = rd.pod<uint64_t>();
    sec.nodes4 = rd.section<QNode4>();
    sec.nodes8 = rd.section<QNode8>();
    sec.prims = rd.section<PrimRec>();
    sec.prim_ng = rd.section<float4>();
    sec.aux = rd.section<PrimAux>();
    sec.materials = rd.section<DevMaterial>();
    sec.uvs = rd.section<float4>();
    sec.tangents = rd.section<float4>();
    sec.texels = rd.section<float4>();
    sec.images = rd.section<DevImage>();
    sec.textures = rd.section<DevTexture>();
    pc.order = rd.vec<uint32_t>();
    pc.tri_to_prim = rd.vec<uint32_t>();
    pc.lights.radius2 = rd.pod<double>();
    pc.lights.punctual_weight = rd.vec<double>();
    pc.lights.emissive = rd.vec<uint32_t>();
    pc.lights.emissive_weight = rd.vec<double>();
    pc.lights.punctual = rd.vec<DevPunctual>();
    if (rd.pos != rd.size) throw std::runtime_error(label + ": inconsistent section sizes");
    return pc;
}

template <class T>
T record(const Section& s, size_t i) {
    T v;
    std::memcpy(&v, static_cast<const uint8_t*>(s.p) + i * sizeof(T), sizeof(T));
    return v;
}

// Every index a kernel or the host follows, checked before anything is
// uploaded, so that a file whose checksum matches but whose contents do not
// describe a scene (edited by hand, or written with another meaning of a
// field) is refused instead of being read out of bounds:
//   section sizes agree with the triangle and primitive counts;
//   wide nodes: inner children have larger indices than their parent (the
//   tree is acyclic, so every walk ends), leaves lie inside the primitives;
//   primitives: triangle ids, materials and flags; the leaf-order maps;
//   materials: texture slots and UV sets, with UVs (and tangents for normal
//   maps) present; textures, images and texel ranges; emissive triangles and
//   punctual light types.
void validate_cache(const ParsedCache& pc, const std::string& label) {
    const SceneSections& sec = pc.sec;
    const uint64_t n = pc.triangles, np = sec.prims.count, nn = sec.nodes4.count, nn8 = sec.nodes8.count;
    auto fail = [&](const std::string& what) { throw std::runtime_error(label + ": " + what); };
    if (np < n || sec.prim_ng.count != np || sec.aux.count != np || pc.order.size() != np || pc.tri_to_prim.size() != n ||
        (sec.uvs.count && sec.uvs.count != 3 * np) || (sec.tangents.count && sec.tangents.count != 3 * np) ||
        (n && (!nn || !nn8)) || (!nn != !nn8) || pc.lights.emissive.size() != pc.lights.emissive_weight.size() ||
        pc.lights.punctual.size() != pc.lights.punctual_weight.size())
        fail("inconsistent section sizes");
    if (np > kLeafMaxFirst + 1 || nn >= kEmptyChild || nn8 >= kEmptyChild || n >= kNoTriangle)
        fail("more records than the device format holds");
    std::atomic<bool> bad_node{false}, bad_prim{false}, bad_map{false};
    auto check_tree = [&](const Section& nodes, auto tag) {
        using Node = decltype(tag);
        constexpr uint32_t width = sizeof(Node{}.child) / sizeof(uint32_t);
        const uint64_t count = nodes.count;
        parallel_chunks(count, size_t(1) << 16, [&](size_t b, size_t e, size_t) {
            for (size_t i = b; i < e && !bad_node; ++i) {
                const Node w = record<Node>(nodes, i);
                const uint32_t kids = w.exps >> 24;
                for (uint32_t e8 = 0; e8 < 3; ++e8) {
                    const uint32_t ex = (w.exps >> (8 * e8)) & 255u;
                    if (ex < 1 || ex > 254) bad_node = true;
                }
                if (kids > width) bad_node = true;
                for (uint32_t k = 0; k < width; ++k) {
                    const uint32_t c = w.child[k];
                    if (c == kEmptyChild) {
                        if (k < kids) bad_node = true;
                        continue;
                    }
                    if (k >= kids ||
                        (ref_is_leaf(c) ? uint64_t(leaf_first(c)) + leaf_count(c) > np : (c <= i || c >= count)))
                        bad_node = true;
                }
            }
        });
    };
    check_tree(sec.nodes4, QNode4{});
    check_tree(sec.nodes8, QNode8{});
    if (bad_node) fail("BVH node refers outside the node or primitive arrays");
    const uint32_t nm = uint32_t(sec.materials.count);
    constexpr uint32_t known_flags = kPrimCull | kPrimNeverHit | kPrimMaskTexture | (3u << kPrimClassShift);
    parallel_chunks(np, size_t(1) << 16, [&](size_t b, size_t e, size_t) {
        for (size_t i = b; i < e; ++i) {
            const PrimRec p = record<PrimRec>(sec.prims, i);
            if (p.id >= n || p.material >= nm || (p.flags & ~known_flags) || pc.order[i] != p.id) bad_prim = true;
        }
    });
    if (bad_prim) fail("primitive record refers outside the triangle or material arrays");
    parallel_chunks(n, size_t(1) << 16, [&](size_t b, size_t e, size_t) {
        for (size_t t = b; t < e; ++t) {
            const uint32_t p = pc.tri_to_prim[t];
            if (p >= np || pc.order[p] != t) bad_map = true;
        }
    });
    if (bad_map) fail("triangle to primitive map inconsistent");
    const uint64_t nt = sec.texels.count, ni = sec.images.count, nx = sec.textures.count;
    for (size_t i = 0; i < ni; ++i) {
        const DevImage im = record<DevImage>(sec.images, i);
        const uint64_t size = uint64_t(uint32_t(im.width)) * uint64_t(uint32_t(im.height));
        if (im.width <= 0 || im.height <= 0 || im.offset > nt || size > nt - im.offset || im.srgb_offset > nt ||
            size > nt - im.srgb_offset)
            fail("image " + std::to_string(i) + " lies outside the texel array");
    }
    for (size_t i = 0; i < nx; ++i) {
        const DevTexture t = record<DevTexture>(sec.textures, i);
        if (t.image < 0 || uint64_t(t.image) >= ni || t.wrap_s < 0 || t.wrap_s > 2 || t.wrap_t < 0 || t.wrap_t > 2)
            fail("texture " + std::to_string(i) + " refers outside the image array");
    }
    for (size_t i = 0; i < nm; ++i) {
        const DevMaterial m = record<DevMaterial>(sec.materials, i);
        bool any = false;
        for (const DevSlot* sl : {&m.base_tex, &m.mr_tex, &m.normal_tex, &m.emissive_tex, &m.transmission_tex}) {
            if (sl->texture < 0) continue;
            any = true;
            if (uint64_t(sl->texture) >= nx || sl->texcoord < 0 || sl->texcoord > 1)
                fail("material " + std::to_string(i) + " refers outside the texture array");
        }
        if ((any || m.maps) && sec.uvs.count != 3 * np) fail("material " + std::to_string(i) + " samples textures but the file holds no UVs");
        if (m.normal_tex.texture >= 0 && sec.tangents.count != 3 * np)
            fail("material " + std::to_string(i) + " has a normal map but the file holds no tangents");
    }
    for (uint32_t t : pc.lights.emissive)
        if (t >= n) fail("emissive triangle outside the triangle array");
    for (const DevPunctual& l : pc.lights.punctual)
        if (l.type > 2) fail("unknown punctual light type");
}

// The sampling-angle table of `device`: built on the host once per process and
// uploaded once per device, then kept until the process ends (128 MiB), shared
// by every context on that device.
void ensure_turn_table(int device) {
    static std::mutex mu;
    static std::map<int, void*> tables;
    std::lock_guard<std::mutex> lock(mu);
    if (tables.count(device)) return;
    const size_t bytes = size_t(kTurnEntries) * sizeof(TurnCS);
    std::vector<TurnCS> host(kTurnEntries);
    fill_turn_table(host.data());
    void* p = nullptr;
    check(hipMalloc(&p, bytes), "turn table");
    check(hipMemcpy(p, host.data(), bytes, hipMemcpyHostToDevice), "turn table upload");
    check(set_turn_table(static_cast<const TurnCS*>(p)), "turn table symbol");
    tables[device] = p;
}

// Registers a resident scene; the caller holds ctx->mu.
uint64_t add_scene(crg_ctx* ctx, std::shared_ptr<SceneRes> res) {
    ctx->engine->add_resident(int64_t(res->bytes));
    const uint64_t id = ctx->next_id++;
    ctx->scenes[id] = std::move(res);
    return id;
}

}  // namespace

extern "C" {

const char* crg_version(void) { return "crgpu 0.2 gfx950 wavefront (binary64-exact predicates, no FP64 instructions)"; }

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
    try {
        check(hipSetDevice(device), "hipSetDevice");
        ensure_turn_table(device);
        check(noalloc_scratch(device, &ctx->scratch), "scratch symbol address");
        check(hipStreamCreateWithFlags(&ctx->stream, hipStreamNonBlocking), "hipStreamCreate");
        ctx->engine = std::make_unique<Engine>(device, unsigned(prop.multiProcessorCount));
    } catch (const std::exception&) {
        if (ctx->stream) (void)hipStreamDestroy(ctx->stream);
        delete ctx;
        return CRG_EDEVICE;
    }
    *out = ctx;
    return CRG_OK;
}

void crg_context_destroy(crg_ctx* ctx) {
    if (!ctx) return;
    (void)hipSetDevice(ctx->device);
    ctx->engine.reset();  // finishes admitted requests, fails queued ones
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
        // Host preparation (finalize, SAH build, flattening) runs without the
        // context lock, so it overlaps renders in flight.
        const auto t0 = std::chrono::steady_clock::now();
        HostScene hs = finalize_scene(*desc);
        SceneArrays arr = scene_arrays(hs);
        SceneLights lights = scene_lights(hs);
        const auto t1 = std::chrono::steady_clock::now();
        std::lock_guard<std::mutex> lock(ctx->mu);
        auto res = make_resident(ctx, sections_of(arr), std::move(arr.bvh.order), std::move(arr.tri_to_prim),
                                 std::move(lights), hs.tris.size());
        const auto t2 = std::chrono::steady_clock::now();
        if (info) {
            *info = scene_info(*res);
            info->build_seconds = std::chrono::duration<double>(t1 - t0).count();
            info->upload_seconds = std::chrono::duration<double>(t2 - t1).count();
        }
        *scene_id = add_scene(ctx, std::move(res));
    });
}

int crg_scene_save(crg_ctx* ctx, uint64_t scene_id, const char* path) {
    return guarded(ctx, [&] {
        if (!path) throw std::invalid_argument("null path");
        std::shared_ptr<SceneRes> r;
        CacheContent c;
        {
            std::lock_guard<std::mutex> lock(ctx->mu);
            auto it = ctx->scenes.find(scene_id);
            if (it == ctx->scenes.end()) throw std::out_of_range("unknown scene handle");
            r = it->second;
            check(hipSetDevice(ctx->device), "hipSetDevice");
            hipStream_t st = ctx->stream;
            const uint64_t np = r->host.prim_count;
            c.nodes4 = download<QNode4>(r->nodes4, r->host.node_count4, st);
            c.nodes8 = download<QNode8>(r->nodes8, r->host.node_count8, st);
            c.prims = download<PrimRec>(r->prims, np, st);
            c.prim_ng = download<float4>(r->prim_ng, np, st);
            c.aux = download<PrimAux>(r->aux, np, st);
            c.materials = download<DevMaterial>(r->materials, r->material_count, st);
            c.uvs = download<float4>(r->uvs, r->host.uvs ? 3 * np : 0, st);
            c.tangents = download<float4>(r->tangents, r->host.tangents ? 3 * np : 0, st);
            c.texels = download<float4>(r->texels, r->texel_count, st);
            c.images = download<DevImage>(r->images, r->image_count, st);
            c.textures = download<DevTexture>(r->textures, r->texture_count, st);
            check(hipStreamSynchronize(st), "scene download");
        }
        write_cache_file(path, r->triangles, r->binary_nodes, sections_of(c), r->prim_to_tri, r->tri_to_prim,
                         r->lights);
    });
}

int crg_scene_load(crg_ctx* ctx, const char* path, uint64_t* scene_id, crg_scene_info* info) {
    return guarded(ctx, [&] {
        if (!path || !scene_id) throw std::invalid_argument("null path or scene id");
        const auto t0 = std::chrono::steady_clock::now();
        const FileBytes file = read_file(path, cache_label(path));
        ParsedCache pc = parse_cache(file, cache_label(path));
        validate_cache(pc, cache_label(path));
        check(hipSetDevice(ctx->device), "hipSetDevice");
        const auto t1 = std::chrono::steady_clock::now();
        std::lock_guard<std::mutex> lock(ctx->mu);
        auto res = make_resident(ctx, pc.sec, std::move(pc.order), std::move(pc.tri_to_prim), std::move(pc.lights),
                                 pc.triangles);
        const auto t2 = std::chrono::steady_clock::now();
        if (info) {
            *info = scene_info(*res);
            info->build_seconds = std::chrono::duration<double>(t1 - t0).count();
            info->upload_seconds = std::chrono::duration<double>(t2 - t1).count();
        }
        *scene_id = add_scene(ctx, std::move(res));
    });
}

int crg_scene_write_cache(const crg_scene_desc* desc, const char* path, crg_scene_info* info, char* message,
                          size_t capacity) {
    return guarded_msg(message, capacity, [&] {
        if (!desc || !path) throw std::invalid_argument("null scene descriptor or path");
        if (desc->triangle_count && (!desc->positions || !desc->material_ids))
            throw std::invalid_argument("scene descriptor without positions or material ids");
        const auto t0 = std::chrono::steady_clock::now();
        HostScene hs = finalize_scene(*desc);
        SceneArrays arr = scene_arrays(hs);
        SceneLights lights = scene_lights(hs);
        const auto t1 = std::chrono::steady_clock::now();
        write_cache_file(path, hs.tris.size(), arr.bvh.binary_nodes, sections_of(arr), arr.bvh.order, arr.tri_to_prim,
                         lights);
        if (info) {
            *info = crg_scene_info{};
            info->triangles = hs.tris.size();
            info->bvh_nodes = arr.bvh.binary_nodes;
            info->wide_nodes = arr.bvh.q4.size();
            info->build_seconds = std::chrono::duration<double>(t1 - t0).count();
        }
    });
}

int crg_scene_check(const char* path, char* message, size_t capacity) {
    return guarded_msg(message, capacity, [&] {
        if (!path) throw std::invalid_argument("null path");
        const FileBytes file = read_file(path, cache_label(path));
        validate_cache(parse_cache(file, cache_label(path)), cache_label(path));
    });
}

int crg_scene_release(crg_ctx* ctx, uint64_t scene_id) {
    return guarded(ctx, [&] {
        std::lock_guard<std::mutex> lock(ctx->mu);
        auto it = ctx->scenes.find(scene_id);
        if (it == ctx->scenes.end()) throw std::out_of_range("unknown scene handle");
        // Requests in flight keep their own reference; memory goes with the last.
        ctx->engine->add_resident(-int64_t(it->second->bytes));
        ctx->scenes.erase(it);
    });
}

int crg_env_upload(crg_ctx* ctx, const crg_env_desc* desc, uint64_t* env_id, uint64_t* device_bytes) {
    return guarded(ctx, [&] {
        if (!desc || !env_id) throw std::invalid_argument("null environment descriptor");
        check(hipSetDevice(ctx->device), "hipSetDevice");
        HostEnv he = build_env(*desc);
        auto res = std::make_shared<EnvRes>();
        std::lock_guard<std::mutex> lock(ctx->mu);
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
        ctx->engine->add_resident(int64_t(res->bytes));
        const uint64_t id = ctx->next_id++;
        ctx->envs[id] = std::move(res);
        *env_id = id;
    });
}

int crg_env_release(crg_ctx* ctx, uint64_t env_id) {
    return guarded(ctx, [&] {
        std::lock_guard<std::mutex> lock(ctx->mu);
        auto it = ctx->envs.find(env_id);
        if (it == ctx->envs.end()) throw std::out_of_range("unknown environment handle");
        ctx->engine->add_resident(-int64_t(it->second->bytes));
        ctx->envs.erase(it);
    });
}

int crg_render_window(crg_ctx* ctx, uint64_t scene_id, uint64_t env_id, const crg_camera* camera,
                      const crg_options* options, int32_t x0, int32_t y0, int32_t width, int32_t height, float* film_rgb,
                      float* samples_rgb, crg_stats* stats) {
    int rc = CRG_OK;
    const int g = guarded(ctx, [&] {
        if (!camera || !options) throw std::invalid_argument("null render argument");
        validate_options(*options);
        if (!film_rgb) throw std::invalid_argument("null film buffer");
        if (x0 < 0 || y0 < 0 || width < 1 || height < 1 || int64_t(x0) + width > options->width ||
            int64_t(y0) + height > options->height)
            throw std::invalid_argument("Render window must lie inside the image");
        crg_request r{};
        r.scene = scene_id;
        r.env = env_id;
        r.camera = *camera;
        r.options = *options;
        r.film_rgb = film_rgb;
        Request q = make_request(ctx, r);
        q.win[0] = x0;
        q.win[1] = y0;
        q.win[2] = width;
        q.win[3] = height;
        q.samples = samples_rgb;
        const uint64_t t = ctx->engine->submit(std::move(q));
        rc = status_of_wait(*ctx->engine, t, -1.0, stats);
    });
    return g != CRG_OK ? g : rc;
}

int crg_device_probe(crg_ctx* ctx, int32_t what, uint64_t scene_id, uint64_t env_id, const crg_camera* camera,
                     int32_t width, int32_t height, const void* in, uint64_t count, void* out) {
    return guarded(ctx, [&] {
        static const size_t in_size[8] = {2 * sizeof(float), sizeof(uint64_t), 3 * sizeof(float), 3 * sizeof(float),
                                          sizeof(crg_probe_bsdf_in), sizeof(crg_probe_bsdf_in),
                                          sizeof(crg_probe_offset_in), sizeof(crg_probe_surface_in)};
        static const size_t out_size[8] = {sizeof(crg_probe_ray), sizeof(crg_probe_env_sample), sizeof(float),
                                           3 * sizeof(float), sizeof(crg_probe_bsdf_sample), 4 * sizeof(float),
                                           sizeof(crg_probe_ray), sizeof(crg_surface)};
        if (what < 0 || what > CRG_PROBE_SURFACE_AT) throw std::invalid_argument("unknown probe function");
        if (count && (!in || !out)) throw std::invalid_argument("null probe records");
        ProbeArgs a{};
        a.what = what;
        a.n = count;
        std::shared_ptr<SceneRes> scene;
        std::shared_ptr<EnvRes> env;
        {
            std::lock_guard<std::mutex> lock(ctx->mu);
            if (what == CRG_PROBE_SURFACE_AT) {
                auto it = ctx->scenes.find(scene_id);
                if (it == ctx->scenes.end()) throw std::out_of_range("unknown scene handle");
                scene = it->second;
            }
            if (what >= CRG_PROBE_ENV_SAMPLE && what <= CRG_PROBE_ENV_RADIANCE) {
                auto it = ctx->envs.find(env_id);
                if (it == ctx->envs.end()) throw std::out_of_range("unknown environment handle");
                env = it->second;
            }
        }
        if (what == CRG_PROBE_CAMERA) {
            if (!camera) throw std::invalid_argument("null camera");
            a.camera = prepare_camera(*camera, width, height);
        }
        std::vector<float> ung;
        std::vector<uint32_t> tri_to_prim_check;
        if (what == CRG_PROBE_OFFSET_RAY) {
            ung.resize(3 * count);
            const auto* q = static_cast<const crg_probe_offset_in*>(in);
            for (uint64_t i = 0; i < count; ++i) unit_vector(q[i].geometric_normal, &ung[3 * i]);
        }
        if (what == CRG_PROBE_SURFACE_AT) {
            const auto* q = static_cast<const crg_probe_surface_in*>(in);
            for (uint64_t i = 0; i < count; ++i)
                if (q[i].hit.triangle >= scene->tri_to_prim.size()) throw std::invalid_argument("hit triangle outside the scene");
        }
        std::lock_guard<std::mutex> lock(ctx->mu);
        check(hipSetDevice(ctx->device), "hipSetDevice");
        hipStream_t st = ctx->stream;
        DevBuf din, dout, dmap, dung;
        din.alloc(count * in_size[what]);
        dout.alloc(count * out_size[what]);
        if (count) check(hipMemcpyAsync(din.p, in, count * in_size[what], hipMemcpyHostToDevice, st), "copy");
        if (scene) {
            upload(dmap, scene->tri_to_prim, st);
            a.scene = scene->desc.as<DevScene>();
            a.tri_to_prim = dmap.as<uint32_t>();
        }
        if (env) a.env = env->desc.as<DevEnv>();
        if (!ung.empty()) {
            upload(dung, ung, st);
            a.ung = dung.as<float>();
        }
        a.in = din.p;
        a.out = dout.p;
        check(launch_probe(a, st), "probe launch");
        if (count) check(hipMemcpyAsync(out, dout.p, count * out_size[what], hipMemcpyDeviceToHost, st), "copy");
        check(hipStreamSynchronize(st), "probe");
    });
}

int crg_trace_rays(crg_ctx* ctx, uint64_t scene_id, const crg_ray* rays, uint64_t count, crg_hit* hits, int any) {
    return guarded(ctx, [&] {
        std::lock_guard<std::mutex> lock(ctx->mu);
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
                                any, ctx->ctl.as<Control>(), int(ctx->engine->tuning().bvh_width), st),
              "trace launch");
        if (count) check(hipMemcpyAsync(hits, ctx->hits.p, count * sizeof(crg_hit), hipMemcpyDeviceToHost, st), "copy");
        check(hipStreamSynchronize(st), "trace");
        for (uint64_t i = 0; i < count; ++i) {
            if (hits[i].hit == 2u) throw std::logic_error("BVH traversal exceeded its bounded stack");
            if (hits[i].hit == 3u) throw std::runtime_error("Non-finite texture coordinates or transform");
        }
    });
}

int crg_trace_rays_noalloc(crg_ctx* ctx, uint64_t scene_id, const crg_ray* rays, uint64_t count, crg_hit* hits,
                           int any) {
    return guarded(ctx, [&] {
        if (count > kNoAllocMaxRays)
            throw std::invalid_argument("crg_trace_rays_noalloc: " + std::to_string(count) + " rays exceed the " +
                                        std::to_string(kNoAllocMaxRays) + "-ray scratch; split the batch");
        if (count && (!rays || !hits)) throw std::invalid_argument("null ray or hit array");
        std::lock_guard<std::mutex> lock(ctx->mu);
        auto it = ctx->scenes.find(scene_id);
        if (it == ctx->scenes.end()) throw std::out_of_range("unknown scene handle");
        check(hipSetDevice(ctx->device), "hipSetDevice");
        hipStream_t st = ctx->stream;
        uint8_t* base = static_cast<uint8_t*>(ctx->scratch);
        Control* ctl = reinterpret_cast<Control*>(base);
        RayIn* drays = reinterpret_cast<RayIn*>(base + kNoAllocCtlBytes);
        HitOut* dhits = reinterpret_cast<HitOut*>(base + kNoAllocCtlBytes + kNoAllocMaxRays * sizeof(RayIn));
        // The control block's counters start at zero on every call (the array
        // keeps the previous call's contents).
        check(launch_clear_words(reinterpret_cast<uint32_t*>(ctl), uint32_t(sizeof(Control) / 4), st), "clear");
        if (count) check(hipMemcpyAsync(drays, rays, count * sizeof(crg_ray), hipMemcpyHostToDevice, st), "copy");
        check(launch_trace_rays(it->second->desc.as<DevScene>(), drays, count, dhits, any, ctl,
                                int(ctx->engine->tuning().bvh_width), st),
              "trace launch");
        if (count) check(hipMemcpyAsync(hits, dhits, count * sizeof(crg_hit), hipMemcpyDeviceToHost, st), "copy");
        check(hipStreamSynchronize(st), "trace");
        for (uint64_t i = 0; i < count; ++i) {
            if (hits[i].hit == 2u) throw std::logic_error("BVH traversal exceeded its bounded stack");
            if (hits[i].hit == 3u) throw std::runtime_error("Non-finite texture coordinates or transform");
        }
    });
}

uint64_t crg_film_hdr_bound(int32_t width, int32_t height) {
    if (width <= 0 || height <= 0) return 0;
    const uint64_t w = uint64_t(width), h = uint64_t(height);
    return cr::hdr_header(width, height).size() + h * (4 + 4 * std::max<uint64_t>(cr::rle_bound(uint32_t(w)), w));
}

namespace {
const char* kPixelsMessage = ": pixels must contain finite, nonnegative linear radiance";

void film_encode(crg_ctx* ctx, const float* film_rgb, int32_t width, int32_t height, float exposure, uint8_t* rgb8,
                 std::vector<uint8_t>* hdr) {
    if (!film_rgb) throw std::invalid_argument("null film");
    if (rgb8) cr::check_dimensions(width, height, "PNG output");
    if (hdr) cr::check_dimensions(width, height, "HDR output");
    if (!rgb8 && !hdr) return;
    const uint32_t w = uint32_t(width), h = uint32_t(height);
    const uint64_t npix = uint64_t(w) * h;
    const bool rle = w >= 8 && w < 32768;
    std::lock_guard<std::mutex> lock(ctx->mu);
    check(hipSetDevice(ctx->device), "hipSetDevice");
    hipStream_t st = ctx->stream;
    const bool exposure_ok = std::isfinite(exposure) && exposure >= -100 && exposure <= 100;
    if (rgb8 && exposure_ok && !(exposure == ctx->cached_exposure)) {
        cr::display_thresholds(std::exp2(double(exposure)), ctx->cached_thresholds);
        ctx->cached_exposure = exposure;
    }
    ctx->film.alloc(npix * 12);
    ctx->first_bad.alloc(8);
    ctx->thresholds.alloc(sizeof(ctx->cached_thresholds));
    if (rgb8) ctx->ldr.alloc(npix * 3);
    if (hdr) ctx->planes.alloc(npix * 4);
    check(hipMemcpyAsync(ctx->film.p, film_rgb, npix * 12, hipMemcpyHostToDevice, st), "film upload");
    check(hipMemsetAsync(ctx->first_bad.p, 0xff, 8, st), "memset");
    if (rgb8)
        check(hipMemcpyAsync(ctx->thresholds.p, ctx->cached_thresholds, sizeof(ctx->cached_thresholds),
                             hipMemcpyHostToDevice, st),
              "thresholds");
    check(cr::launch_film_encode(ctx->film.as<float>(), npix, w, ctx->thresholds.as<float>(),
                                 rgb8 ? ctx->ldr.as<uint8_t>() : nullptr, hdr ? ctx->planes.as<uint8_t>() : nullptr,
                                 ctx->first_bad.as<uint32_t>(), st),
          "film encode launch");
    const uint32_t bound = cr::rle_bound(w);
    if (hdr && rle) {
        ctx->streams.alloc(uint64_t(h) * 4 * bound);
        ctx->lens.alloc(uint64_t(h) * 4 * sizeof(uint32_t));
        check(cr::launch_film_rle(ctx->planes.as<uint8_t>(), w, h, ctx->streams.as<uint8_t>(), ctx->lens.as<uint32_t>(),
                                  st),
              "film rle launch");
    }
    uint32_t bad[2];
    check(hipMemcpyAsync(bad, ctx->first_bad.p, 8, hipMemcpyDeviceToHost, st), "copy");
    check(hipStreamSynchronize(st), "film encode");
    if (rgb8) {
        if (bad[0] != ~0u) throw std::runtime_error(std::string("PNG output") + kPixelsMessage);
        cr::check_exposure(exposure);
        check(hipMemcpyAsync(rgb8, ctx->ldr.p, npix * 3, hipMemcpyDeviceToHost, st), "copy");
    }
    if (hdr) {
        if (std::min(bad[0], bad[1]) != ~0u) {
            if (bad[0] <= bad[1]) throw std::runtime_error(std::string("HDR output") + kPixelsMessage);
            throw std::runtime_error("HDR output: radiance exceeds the Radiance RGBE exponent range");
        }
        const std::string head = cr::hdr_header(width, height);
        hdr->assign(head.begin(), head.end());
        if (rle) {
            std::vector<uint32_t> lens(size_t(h) * 4);
            std::vector<uint8_t> streams(size_t(h) * 4 * bound);
            check(hipMemcpyAsync(lens.data(), ctx->lens.p, lens.size() * 4, hipMemcpyDeviceToHost, st), "copy");
            check(hipMemcpyAsync(streams.data(), ctx->streams.p, streams.size(), hipMemcpyDeviceToHost, st), "copy");
            check(hipStreamSynchronize(st), "film copy");
            size_t total = hdr->size();
            for (uint32_t v : lens) total += v;
            hdr->reserve(total + size_t(h) * 4);
            for (uint32_t y = 0; y < h; ++y) {
                const uint8_t sh[4] = {2, 2, uint8_t((w >> 8) & 0xff), uint8_t(w & 0xff)};
                hdr->insert(hdr->end(), sh, sh + 4);
                for (uint32_t c = 0; c < 4; ++c) {
                    const size_t j = size_t(y) * 4 + c;
                    if (lens[j] > bound) throw std::logic_error("run-length code exceeds its bound");
                    const uint8_t* p = streams.data() + j * bound;
                    hdr->insert(hdr->end(), p, p + lens[j]);
                }
            }
        } else {  // flat RGBE quadruples
            std::vector<uint8_t> planes(npix * 4);
            check(hipMemcpyAsync(planes.data(), ctx->planes.p, planes.size(), hipMemcpyDeviceToHost, st), "copy");
            check(hipStreamSynchronize(st), "film copy");
            const size_t off = hdr->size();
            hdr->resize(off + npix * 4);
            for (uint32_t y = 0; y < h; ++y)
                for (uint32_t x = 0; x < w; ++x)
                    for (uint32_t c = 0; c < 4; ++c)
                        (*hdr)[off + (size_t(y) * w + x) * 4 + c] = planes[size_t(y) * 4 * w + size_t(c) * w + x];
        }
    }
    check(hipStreamSynchronize(st), "film copy");
}
}  // namespace

int crg_film_encode(crg_ctx* ctx, const float* film_rgb, int32_t width, int32_t height, float exposure, uint8_t* rgb8,
                    uint8_t* hdr, uint64_t hdr_capacity, uint64_t* hdr_size) {
    return guarded(ctx, [&] {
        std::vector<uint8_t> file;
        film_encode(ctx, film_rgb, width, height, exposure, rgb8, hdr ? &file : nullptr);
        if (hdr) {
            if (hdr_size) *hdr_size = file.size();
            if (file.size() > hdr_capacity) throw std::invalid_argument("HDR output buffer too small");
            std::memcpy(hdr, file.data(), file.size());
        }
    });
}

int crg_write_png(crg_ctx* ctx, const char* path, const float* film_rgb, int32_t width, int32_t height,
                  float exposure) {
    return guarded(ctx, [&] {
        if (!path) throw std::invalid_argument("null path");
        cr::check_dimensions(width, height, "PNG output");
        std::vector<uint8_t> rgb8(size_t(width) * size_t(height) * 3);
        film_encode(ctx, film_rgb, width, height, exposure, rgb8.data(), nullptr);
        if (!cr::write_png_rgb8(path, width, height, rgb8.data()))
            throw std::runtime_error(std::string("Failed writing image '") + path + "'");
    });
}

int crg_write_hdr(crg_ctx* ctx, const char* path, const float* film_rgb, int32_t width, int32_t height) {
    return guarded(ctx, [&] {
        if (!path) throw std::invalid_argument("null path");
        std::vector<uint8_t> file;
        film_encode(ctx, film_rgb, width, height, 0.0f, nullptr, &file);
        std::FILE* f = std::fopen(path, "wb");
        const bool ok = f && std::fwrite(file.data(), 1, file.size(), f) == file.size();
        if (!f || std::fclose(f) != 0 || !ok) throw std::runtime_error(std::string("Failed writing image '") + path + "'");
    });
}

int crg_render(crg_ctx* ctx, uint64_t scene_id, uint64_t env_id, const crg_camera* camera, const crg_options* options,
               float* film_rgb, crg_stats* stats) {
    int rc = CRG_OK;
    const int g = guarded(ctx, [&] {
        if (!camera || !options) throw std::invalid_argument("null render argument");
        validate_options(*options);
        if (!film_rgb) throw std::invalid_argument("null film buffer");
        crg_request r{};
        r.scene = scene_id;
        r.env = env_id;
        r.camera = *camera;
        r.options = *options;
        r.film_rgb = film_rgb;
        const uint64_t t = ctx->engine->submit(make_request(ctx, r));
        rc = status_of_wait(*ctx->engine, t, -1.0, stats);
    });
    return g != CRG_OK ? g : rc;
}

int crg_render_batch(crg_ctx* ctx, const crg_request* requests, uint32_t count, int* status, crg_stats* stats) {
    int first = CRG_OK;
    std::string first_msg;
    const int g = guarded(ctx, [&] {
        if (count && !requests) throw std::invalid_argument("null request array");
        std::vector<uint64_t> tickets(count, 0);
        std::vector<int> rcs(count, CRG_OK);
        std::vector<std::string> msgs(count);
        for (uint32_t i = 0; i < count; ++i) {
            try {
                tickets[i] = ctx->engine->submit(make_request(ctx, requests[i]));
            } catch (const std::bad_alloc&) {
                rcs[i] = CRG_ENOMEM;
                msgs[i] = "out of memory";
            } catch (const std::out_of_range& e) {
                rcs[i] = CRG_ENOTFOUND;
                msgs[i] = e.what();
            } catch (const std::exception& e) {
                rcs[i] = CRG_EINVAL;
                msgs[i] = e.what();
            }
        }
        for (uint32_t i = 0; i < count; ++i) {
            crg_stats st{};
            if (tickets[i]) rcs[i] = ctx->engine->wait(tickets[i], -1.0, &st, &msgs[i]);
            if (stats) stats[i] = st;
            if (status) status[i] = rcs[i];
            if (rcs[i] != CRG_OK && first == CRG_OK) {
                first = rcs[i];
                first_msg = "request " + std::to_string(i) + ": " + msgs[i];
            }
        }
    });
    if (g != CRG_OK) return g;
    if (first != CRG_OK) {
        std::lock_guard<std::mutex> lock(ctx->mu);
        ctx->error = first_msg;
    }
    return first;
}

int crg_submit(crg_ctx* ctx, const crg_request* request, uint64_t* ticket) {
    return guarded(ctx, [&] {
        if (!request || !ticket) throw std::invalid_argument("null request");
        *ticket = ctx->engine->submit(make_request(ctx, *request));
    });
}

int crg_wait(crg_ctx* ctx, uint64_t ticket, double timeout_seconds, crg_stats* stats) {
    int rc = CRG_OK;
    const int g = guarded(ctx, [&] { rc = status_of_wait(*ctx->engine, ticket, timeout_seconds, stats); });
    return g != CRG_OK ? g : rc;
}

int crg_request_progress(crg_ctx* ctx, uint64_t ticket, uint64_t* complete, uint64_t* total) {
    return guarded(ctx, [&] {
        if (!complete || !total) throw std::invalid_argument("null progress output");
        ctx->engine->progress(ticket, complete, total);
    });
}

int crg_set_memory_budget(crg_ctx* ctx, uint64_t bytes) {
    return guarded(ctx, [&] { ctx->engine->set_budget(bytes); });
}

int crg_engine_info_get(crg_ctx* ctx, crg_engine_info* info) {
    return guarded(ctx, [&] {
        if (!info) throw std::invalid_argument("null info");
        *info = ctx->engine->info();
    });
}

int crg_get_tuning(crg_ctx* ctx, crg_tuning* out) {
    return guarded(ctx, [&] {
        if (!out) throw std::invalid_argument("null tuning");
        *out = ctx->engine->tuning();
    });
}

int crg_set_tuning(crg_ctx* ctx, const crg_tuning* tuning) {
    return guarded(ctx, [&] {
        if (!tuning) throw std::invalid_argument("null tuning");
        ctx->engine->set_tuning(*tuning);
    });
}

int crg_autotune(crg_ctx* ctx, const crg_request* sample, crg_tuning* out, double* samples_per_second) {
    return guarded(ctx, [&] {
        if (!sample) throw std::invalid_argument("null sample request");
        Engine& eng = *ctx->engine;
        const crg_tuning start = eng.tuning();
        const uint64_t items = uint64_t(sample->options.width) * sample->options.height * sample->options.samples;
        auto time_of = [&](const crg_tuning& t) {
            eng.set_tuning(t);
            double best = 1e300;
            for (int rep = 0; rep < 3; ++rep) {  // the first run also absorbs (re)configuration
                const auto t0 = std::chrono::steady_clock::now();
                std::string err;
                const int rc = eng.wait(eng.submit(make_request(ctx, *sample)), -1.0, nullptr, &err);
                if (rc != CRG_OK) throw std::runtime_error("autotune sample render failed: " + err);
                best = std::min(best, std::chrono::duration<double>(std::chrono::steady_clock::now() - t0).count());
            }
            return best;
        };
        crg_tuning best = start;
        double best_s = time_of(best);
        auto try_field = [&](uint32_t crg_tuning::*field, std::initializer_list<uint32_t> values) {
            for (uint32_t v : values) {
                crg_tuning t = best;
                t.*field = v;
                if (std::memcmp(&t, &best, sizeof(t)) == 0) continue;
                if (field == &crg_tuning::pool_paths && uint64_t(v) > 2 * items && v > (1u << 16)) continue;
                const double s = time_of(t);
                if (s < best_s * 0.98) {
                    best_s = s;
                    best = t;
                }
            }
        };
        try_field(&crg_tuning::bvh_width, {4, 8});
        // A sample cannot time pools it does not fill: when it has fewer than half
        // as many work items as the configured pool has slots, the pool is kept (a
        // larger pool only pays on larger work; 64 batched requests measured 446 ->
        // 523 Msamples/s from 4 M to 32 M slots).
        if (uint64_t(best.pool_paths) <= 2 * items)
            try_field(&crg_tuning::pool_paths, {1u << 20, 1u << 21, 1u << 22, 1u << 23, 1u << 24, 1u << 25});
        try_field(&crg_tuning::groups_per_cu, {4, 8, 16});
        try_field(&crg_tuning::workgroup, {64, 128, 256});
        try_field(&crg_tuning::iterations_per_graph, {2, 4, 8});
        eng.set_tuning(best);
        if (out) *out = best;
        if (samples_per_second) *samples_per_second = double(items) / best_s;
    });
}

int crg_set_work_counters(crg_ctx* ctx, int enable) {
    return guarded(ctx, [&] { ctx->engine->set_counting(enable != 0); });
}

int crg_set_profiling(crg_ctx* ctx, int enable) {
    return guarded(ctx, [&] { ctx->engine->set_profiling(enable != 0); });
}

int crg_kernel_times(crg_ctx* ctx, crg_kernel_time* out, int max, int* count) {
    return guarded(ctx, [&] { ctx->engine->kernel_times(out, max, count); });
}

}  // extern "C"
