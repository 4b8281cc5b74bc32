// host_scene.cpp — scene finalization, 4-wide SAH build, light/environment
// inversion tables and camera constants, all on the host.
#include "host_scene.h"

#include <algorithm>
#include <array>
#include <atomic>
#include <cmath>
#include <cstring>
#include <future>
#include <limits>
#include <numeric>
#include <sstream>
#include <thread>

#include "parallel.h"

namespace cr {
namespace {

constexpr float kHostPi = 3.14159265358979323846f;

struct DV {
    double x, y, z;
};
DV dv(HVec3 v) { return {v.x, v.y, v.z}; }
DV operator+(DV a, DV b) { return {a.x + b.x, a.y + b.y, a.z + b.z}; }
DV operator-(DV a, DV b) { return {a.x - b.x, a.y - b.y, a.z - b.z}; }
DV operator*(DV a, double s) { return {a.x * s, a.y * s, a.z * s}; }
double ddot_(DV a, DV b) { return a.x * b.x + a.y * b.y + a.z * b.z; }
DV dcross(DV a, DV b) { return {a.y * b.z - a.z * b.y, a.z * b.x - a.x * b.z, a.x * b.y - a.y * b.x}; }

// unit() of the reference (src/geometry.cpp:31-37).
HVec3 unit_of(DV v, HVec3 fallback = {0, 0, 0}) {
    const double sq = ddot_(v, v);
    if (!(sq > 0) || !std::isfinite(sq)) return fallback;
    const double inv = 1 / std::sqrt(sq);
    return {float(v.x * inv), float(v.y * inv), float(v.z * inv)};
}
// orthogonalized() of the reference (src/geometry.cpp:40-47).
HVec3 orthogonal_of(DV v, HVec3 n, HVec3 fallback = {0, 0, 0}) {
    const double nn = ddot_(dv(n), dv(n));
    if (!(nn > 0)) return fallback;
    const DV proj = v - dv(n) * (ddot_(v, dv(n)) / nn);
    constexpr double tol = 4 * std::numeric_limits<float>::epsilon();
    if (ddot_(proj, proj) <= ddot_(v, v) * tol * tol) return fallback;
    return unit_of(proj, fallback);
}
HVec3 hcross(HVec3 a, HVec3 b) { return {a.y * b.z - a.z * b.y, a.z * b.x - a.x * b.z, a.x * b.y - a.y * b.x}; }
bool nonzero3_(HVec3 v) { return v.x != 0 || v.y != 0 || v.z != 0; }

// UV-to-position differential directions (reference src/geometry.cpp:98-116).
bool uv_directions(const HTriangle& t, int set, DV& tangent, DV& bitangent) {
    const double du1 = double(t.uv[1][2 * set]) - t.uv[0][2 * set];
    const double dv1 = double(t.uv[1][2 * set + 1]) - t.uv[0][2 * set + 1];
    const double du2 = double(t.uv[2][2 * set]) - t.uv[0][2 * set];
    const double dv2 = double(t.uv[2][2 * set + 1]) - t.uv[0][2 * set + 1];
    const double det = du1 * dv2 - dv1 * du2;
    if (det == 0) return false;
    const DV e1 = dv(t.p[1]) - dv(t.p[0]), e2 = dv(t.p[2]) - dv(t.p[0]);
    const double sign = det < 0 ? -1 : 1;
    tangent = (e1 * dv2 - e2 * dv1) * sign;
    bitangent = (e2 * du1 - e1 * du2) * sign;
    return nonzero3_(unit_of(tangent)) && nonzero3_(unit_of(bitangent));
}

}  // namespace

void fill_turn_table(TurnCS* out) {
    parallel_chunks(kTurnEntries, 1 << 16, [&](size_t b, size_t e, size_t) {
        for (size_t k = b; k < e; ++k) {
            const float u = float(uint32_t(k)) * (1.0f / 16777216.0f);
            const float phi = 2 * kHostPi * u;
            out[k] = TurnCS{std::cos(phi), std::sin(phi)};
        }
    });
}

namespace {

// sRGB to linear of one channel, the reference's expression (reference src/math.h:61).
float srgb_to_linear(float v) { return v <= 0.04045f ? v / 12.92f : std::pow((v + 0.055f) / 1.055f, 2.4f); }

bool fin(float v) { return std::isfinite(v); }
bool fin3(const float* v) { return fin(v[0]) && fin(v[1]) && fin(v[2]); }
bool nonzero3(HVec3 v) { return v.x != 0 || v.y != 0 || v.z != 0; }
float fdot(HVec3 a, HVec3 b) { return a.x * b.x + a.y * b.y + a.z * b.z; }
bool in_unit(float v) { return std::isfinite(v) && v >= 0 && v <= 1; }
float luminance3(const float* c) { return c[0] * 0.2126f + c[1] * 0.7152f + c[2] * 0.0722f; }

}  // namespace

HostScene finalize_scene(const crg_scene_desc& d) {
    HostScene s;
    s.materials.assign(d.materials, d.materials + d.material_count);
    s.lights.assign(d.lights, d.lights + d.light_count);
    // Images in order: dimensions, then finiteness (the first failing image is
    // reported), copied into texel records in the same parallel pass.
    for (uint32_t i = 0; i < d.image_count; ++i) {
        const crg_image& im = d.images[i];
        if (im.width <= 0 || im.height <= 0 || !im.rgba)
            throw std::runtime_error("Image " + std::to_string(i) + " has invalid dimensions or pixel count");
        HImage h;
        h.width = im.width;
        h.height = im.height;
        h.texels.resize(size_t(im.width) * size_t(im.height));
        std::atomic<bool> finite{true};
        parallel_chunks(h.texels.size(), 65536, [&](size_t b, size_t e, size_t) {
            bool ok = true;
            for (size_t k = b; k < e; ++k) {
                const float* px = im.rgba + 4 * k;
                ok &= fin(px[0]) & fin(px[1]) & fin(px[2]) & fin(px[3]);
                h.texels[k] = make_float4(px[0], px[1], px[2], px[3]);
            }
            if (!ok) finite = false;
        });
        if (!finite) throw std::runtime_error("Image " + std::to_string(i) + " contains non-finite pixels");
        s.images.push_back(std::move(h));
    }
    for (uint32_t i = 0; i < d.texture_count; ++i) {
        const crg_texture& t = d.textures[i];
        if (t.image < 0 || uint32_t(t.image) >= d.image_count)
            throw std::runtime_error("Texture " + std::to_string(i) + " has an invalid image index");
        auto ok = [](int w) { return w == 0 || w == 1 || w == 2; };
        if (!ok(t.wrap_s) || !ok(t.wrap_t))
            throw std::runtime_error("Texture " + std::to_string(i) + " has an invalid wrap mode");
        s.textures.push_back(t);
    }
    for (size_t i = 0; i < s.materials.size(); ++i) {
        const crg_material& m = s.materials[i];
        const std::string label = "Material " + std::to_string(i);
        if (!fin(m.base_color[0]) || !fin(m.base_color[1]) || !fin(m.base_color[2]) || !fin(m.base_color[3]) ||
            !in_unit(m.base_color[0]) || !in_unit(m.base_color[1]) || !in_unit(m.base_color[2]) ||
            !in_unit(m.base_color[3]) || !in_unit(m.metallic) || !in_unit(m.roughness) || !in_unit(m.transmission) ||
            !in_unit(m.alpha_cutoff) || !fin3(m.emission) || m.emission[0] < 0 || m.emission[1] < 0 ||
            m.emission[2] < 0 || !fin(m.ior) || m.ior <= 0 || !fin(m.thickness) || m.thickness < 0 ||
            std::isnan(m.attenuation_distance) || m.attenuation_distance <= 0 || !in_unit(m.attenuation_color[0]) ||
            !in_unit(m.attenuation_color[1]) || !in_unit(m.attenuation_color[2]) || !fin(m.normal_scale))
            throw std::runtime_error(label + " has non-finite or out-of-range parameters");
        for (const crg_texture_slot* slot :
             {&m.base_color_tex, &m.metallic_roughness_tex, &m.normal_tex, &m.emissive_tex, &m.transmission_tex}) {
            if (slot->texture < -1 || (slot->texture >= 0 && uint32_t(slot->texture) >= d.texture_count))
                throw std::runtime_error(label + " has an invalid texture index");
            if (slot->texture >= 0 && (slot->texcoord < 0 || slot->texcoord > 1 || !fin(slot->offset[0]) ||
                                       !fin(slot->offset[1]) || !fin(slot->scale[0]) || !fin(slot->scale[1]) ||
                                       !fin(slot->rotation)))
                throw std::runtime_error(label + " has invalid texture coordinates or transform");
        }
        if (m.normal_tex.texture >= 0 && (m.normal_tex.scale[0] == 0 || m.normal_tex.scale[1] == 0))
            s.warnings.push_back(label + " has a singular normal-map UV transform; using the untransformed tangent frame");
    }
    const uint64_t n = d.triangle_count;
    // Validation and per-triangle preparation run in contiguous chunks on all
    // host threads; the first failing triangle by index decides the error, as
    // in a sequential pass.
    constexpr size_t kMinChunk = 16384;
    const size_t nchunks = chunk_count(n, kMinChunk);
    std::vector<uint64_t> bad(nchunks, UINT64_MAX);
    std::vector<int> bad_kind(nchunks, 0);
    parallel_chunks(n, kMinChunk, [&](size_t b, size_t e, size_t c) {
        for (uint64_t i = b; i < e; ++i) {
            if (d.material_ids[i] >= d.material_count) {
                bad[c] = i, bad_kind[c] = 1;
                return;
            }
            bool ok = true;
            for (int k = 0; k < 9; ++k) ok = ok && fin(d.positions[9 * i + k]);
            if (d.normals)
                for (int k = 0; k < 9; ++k) ok = ok && fin(d.normals[9 * i + k]);
            if (d.tangents)
                for (int k = 0; k < 12; ++k) ok = ok && fin(d.tangents[12 * i + k]);
            if (d.uvs)
                for (int k = 0; k < 12; ++k) ok = ok && fin(d.uvs[12 * i + k]);
            if (!ok) {
                bad[c] = i, bad_kind[c] = 2;
                return;
            }
        }
    });
    for (size_t c = 0; c < nchunks; ++c)
        if (bad[c] != UINT64_MAX)
            throw std::runtime_error("Triangle " + std::to_string(bad[c]) +
                                     (bad_kind[c] == 1 ? " has an invalid material index"
                                                       : " contains non-finite vertex data"));
    struct Chunk {
        std::vector<HTriangle> tris;
        size_t degenerate = 0, first_degenerate = 0, generated = 0, fallback_tangents = 0;
        uint64_t too_large = UINT64_MAX;
    };
    std::vector<Chunk> parts(nchunks);
    parallel_chunks(n, kMinChunk, [&](size_t b, size_t e, size_t c) {
        Chunk& out = parts[c];
        out.tris.reserve(e - b);
        for (uint64_t i = b; i < e; ++i) {
            HTriangle t;
            for (int k = 0; k < 3; ++k) {
                t.p[k] = {d.positions[9 * i + 3 * k], d.positions[9 * i + 3 * k + 1], d.positions[9 * i + 3 * k + 2]};
                t.n[k] = d.normals ? HVec3{d.normals[9 * i + 3 * k], d.normals[9 * i + 3 * k + 1],
                                           d.normals[9 * i + 3 * k + 2]}
                                   : HVec3{0, 0, 0};
                for (int q = 0; q < 4; ++q) {
                    t.tan[k][q] = d.tangents ? d.tangents[12 * i + 4 * k + q] : (q == 3 ? 1.0f : 0.0f);
                    t.uv[k][q] = d.uvs ? d.uvs[12 * i + 4 * k + q] : 0.0f;
                }
            }
            t.material = d.material_ids[i];
            const DV normal = dcross(dv(t.p[1]) - dv(t.p[0]), dv(t.p[2]) - dv(t.p[0]));
            const double twice = std::sqrt(ddot_(normal, normal));
            if (twice * 0.5 > std::numeric_limits<float>::max()) {
                out.too_large = i;
                return;
            }
            t.area = float(twice * 0.5);
            if (!(t.area > 0)) {
                if (out.degenerate == 0) out.first_degenerate = i;
                ++out.degenerate;
                continue;
            }
            t.ng = unit_of(normal);
            // Tangent frames (reference src/geometry.cpp:365-388): vertex tangents
            // orthogonalized against the repaired normal, else UV-derived in the
            // normal map's UV set, else an orthogonal basis.
            const crg_material& mat = s.materials[t.material];
            const int set = mat.normal_tex.texture >= 0 ? mat.normal_tex.texcoord : 0;
            DV ut{0, 0, 0}, ub{0, 0, 0};
            const bool usable = uv_directions(t, set, ut, ub);
            for (int k = 0; k < 3; ++k) {
                HVec3 nn = unit_of(dv(t.n[k]));
                float hand = t.tan[k][3] < 0 ? -1.0f : 1.0f;
                if (!nonzero3(nn) || fdot(nn, t.ng) == 0) {
                    nn = t.ng;
                    ++out.generated;
                } else if (fdot(nn, t.ng) < 0) {
                    nn = {-nn.x, -nn.y, -nn.z};
                    hand = -hand;
                }
                HVec3 tg = orthogonal_of(DV{t.tan[k][0], t.tan[k][1], t.tan[k][2]}, nn);
                if (!nonzero3(tg)) {
                    if (usable) tg = orthogonal_of(ut, nn);
                    if (nonzero3(tg)) {
                        const HVec3 bt = hcross(nn, tg);
                        hand = ddot_(dv(bt), ub) < 0 ? -1.0f : 1.0f;
                    } else {
                        const V3 fx = make_frame(V3{nn.x, nn.y, nn.z}).x;
                        tg = {fx.x, fx.y, fx.z};
                        hand = 1;
                        if (mat.normal_tex.texture >= 0) ++out.fallback_tangents;
                    }
                }
                t.n[k] = nn;
                t.tan[k][0] = tg.x;
                t.tan[k][1] = tg.y;
                t.tan[k][2] = tg.z;
                t.tan[k][3] = hand;
            }
            t.ung = unit_of(dv(t.ng));
            out.tris.push_back(t);
        }
    });
    size_t degenerate = 0, first_degenerate = 0, generated = 0, fallback_tangents = 0, total = 0;
    for (const Chunk& c : parts) {
        if (c.too_large != UINT64_MAX)
            throw std::runtime_error("Triangle " + std::to_string(c.too_large) +
                                     " area exceeds the supported finite range");
        if (c.degenerate && degenerate == 0) first_degenerate = c.first_degenerate;
        degenerate += c.degenerate;
        generated += c.generated;
        fallback_tangents += c.fallback_tangents;
        total += c.tris.size();
    }
    s.tris.reserve(total);
    for (Chunk& c : parts) s.tris.insert(s.tris.end(), c.tris.begin(), c.tris.end());
    if (degenerate)
        s.warnings.push_back("Skipped " + std::to_string(degenerate) + " degenerate triangles (first input triangle " +
                             std::to_string(first_degenerate) + ")");
    if (generated)
        s.warnings.push_back("Generated flat normals for " + std::to_string(generated) +
                             " missing or tangential vertex normals");
    if (fallback_tangents)
        s.warnings.push_back("Used orthogonal fallback tangents for " + std::to_string(fallback_tangents) +
                             " normal-mapped vertices with degenerate UVs");
    return s;
}

// ------------------------------------------------------------------ BVH
namespace {

// Build record: bounds, centroid and the finalized triangle index.  The
// builder partitions these records themselves (not an index array), so every
// scan over a range reads memory sequentially.
struct BPrim {
    float lo[3], hi[3], c[3];
    uint32_t id;
};

struct Box {
    float lo[3] = {INFINITY, INFINITY, INFINITY}, hi[3] = {-INFINITY, -INFINITY, -INFINITY};
    void grow(const float* l, const float* h) {
        for (int a = 0; a < 3; ++a) {
            lo[a] = std::min(lo[a], l[a]);
            hi[a] = std::max(hi[a], h[a]);
        }
    }
    void grow(const Box& b) { grow(b.lo, b.hi); }
    double area() const {
        if (lo[0] > hi[0] || lo[1] > hi[1] || lo[2] > hi[2]) return 0;
        const double dx = double(hi[0]) - lo[0], dy = double(hi[1]) - lo[1], dz = double(hi[2]) - lo[2];
        return 2 * (dx * dy + dy * dz + dz * dx);
    }
};

constexpr int kBins = 16;
#ifndef CR_LEAF_SIZE
#define CR_LEAF_SIZE 4
#endif
#ifndef CR_SAH_TRAVERSAL
#define CR_SAH_TRAVERSAL 1.0
#endif
constexpr uint32_t kLeafSize = CR_LEAF_SIZE;
constexpr double kSahTraversal = CR_SAH_TRAVERSAL;  // cost of a node step relative to a triangle test
constexpr int kMaxWideDepth = 40;

struct Builder {
    std::vector<BPrim>& prims;
    std::vector<WNode>& nodes;
    std::atomic<uint32_t> next{0};
    std::atomic<uint64_t> leaves{0};

    // Ranges above this size are scanned by all host threads (the top levels
    // of the tree, where the subtree tasks have not fanned out yet).
    static constexpr uint32_t kParallelScan = 1u << 17;

    Box bounds(uint32_t b, uint32_t e, Box* centers) const {
        auto scan = [&](uint32_t lo, uint32_t hi, Box& box, Box& cen) {
            for (uint32_t i = lo; i < hi; ++i) {
                const BPrim& p = prims[i];
                box.grow(p.lo, p.hi);
                cen.grow(p.c, p.c);
            }
        };
        Box box, cen;
        if (e - b < kParallelScan) {
            scan(b, e, box, cen);
        } else {
            const size_t nc = chunk_count(e - b, kParallelScan / 4);
            std::vector<Box> bx(nc), cx(nc);
            parallel_chunks(e - b, kParallelScan / 4,
                            [&](size_t lo, size_t hi, size_t c) { scan(b + uint32_t(lo), b + uint32_t(hi), bx[c], cx[c]); });
            for (size_t c = 0; c < nc; ++c) {
                box.grow(bx[c]);
                cen.grow(cx[c]);
            }
        }
        if (centers) *centers = cen;
        return box;
    }

    // Binned SAH split of [b,e); returns the split point or 0 when a leaf is cheaper.
    uint32_t sah_split(uint32_t b, uint32_t e) {
        Box cen;
        const Box box = bounds(b, e, &cen);
        const double parent = box.area();
        if (!(parent > 0)) return 0;
        struct Bin {
            Box box;
            uint32_t count = 0;
        };
        using Bins = std::array<std::array<Bin, kBins>, 3>;
        Bins bins{};
        double scale[3];
        for (int a = 0; a < 3; ++a) {
            const double ext = double(cen.hi[a]) - cen.lo[a];
            scale[a] = ext > 0 ? kBins / ext : 0.0;
        }
        auto bin_at = [&](float c, int a) {
            return std::clamp(int((double(c) - cen.lo[a]) * scale[a]), 0, kBins - 1);
        };
        auto fill = [&](uint32_t lo, uint32_t hi, Bins& out) {
            for (uint32_t i = lo; i < hi; ++i) {
                const BPrim& p = prims[i];
                for (int a = 0; a < 3; ++a) {
                    Bin& bn = out[a][bin_at(p.c[a], a)];
                    bn.box.grow(p.lo, p.hi);
                    ++bn.count;
                }
            }
        };
        if (e - b < kParallelScan) {
            fill(b, e, bins);
        } else {  // per-thread bins, merged (min/max and counts: order-independent)
            const size_t nc = chunk_count(e - b, kParallelScan / 4);
            std::vector<Bins> part(nc);
            parallel_chunks(e - b, kParallelScan / 4,
                            [&](size_t lo, size_t hi, size_t c) { fill(b + uint32_t(lo), b + uint32_t(hi), part[c]); });
            for (const Bins& pb : part)
                for (int a = 0; a < 3; ++a)
                    for (int k = 0; k < kBins; ++k) {
                        if (pb[a][k].count) bins[a][k].box.grow(pb[a][k].box);
                        bins[a][k].count += pb[a][k].count;
                    }
        }
        int best_axis = -1, best_split = -1;
        double best = double(e - b) * parent;
        for (int a = 0; a < 3; ++a) {
            if (cen.lo[a] == cen.hi[a]) continue;
            std::array<Box, kBins> rb{};
            std::array<uint32_t, kBins> rc{};
            Box r;
            uint32_t cnt = 0;
            for (int k = kBins - 1; k >= 0; --k) {
                if (bins[a][k].count) r.grow(bins[a][k].box);
                cnt += bins[a][k].count;
                rb[k] = r;
                rc[k] = cnt;
            }
            Box l;
            cnt = 0;
            for (int k = 0; k < kBins - 1; ++k) {
                if (bins[a][k].count) l.grow(bins[a][k].box);
                cnt += bins[a][k].count;
                if (!cnt || !rc[k + 1]) continue;
                const double cost = kSahTraversal * parent + cnt * l.area() + rc[k + 1] * rb[k + 1].area();
                if (cost < best) {
                    best = cost;
                    best_axis = a;
                    best_split = k;
                }
            }
        }
        if (best_axis < 0) return 0;
        const int a = best_axis;
        auto left = [&](const BPrim& p) { return bin_at(p.c[a], a) <= best_split; };
        uint32_t m;
        if (e - b < kParallelScan) {
            m = uint32_t(std::partition(prims.data() + b, prims.data() + e, left) - prims.data());
        } else {  // stable partition by chunks: counts, offsets, scatter
            const size_t nc = chunk_count(e - b, kParallelScan / 4);
            std::vector<uint32_t> nleft(nc, 0);
            std::vector<BPrim> tmp(e - b);
            parallel_chunks(e - b, kParallelScan / 4, [&](size_t lo, size_t hi, size_t c) {
                uint32_t k = 0;
                for (size_t i = lo; i < hi; ++i) k += left(prims[b + i]) ? 1u : 0u;
                nleft[c] = k;
            });
            std::vector<uint32_t> loff(nc), roff(nc);
            uint32_t total_left = 0;
            for (size_t c = 0; c < nc; ++c) total_left += nleft[c];
            uint32_t l = 0, r = total_left;
            for (size_t c = 0; c < nc; ++c) {
                loff[c] = l;
                roff[c] = r;
                l += nleft[c];
                r += uint32_t((c + 1) * (e - b) / nc - c * (e - b) / nc) - nleft[c];
            }
            parallel_chunks(e - b, kParallelScan / 4, [&](size_t lo, size_t hi, size_t c) {
                uint32_t li = loff[c], ri = roff[c];
                for (size_t i = lo; i < hi; ++i) {
                    const BPrim& v = prims[b + i];
                    if (left(v)) tmp[li++] = v;
                    else tmp[ri++] = v;
                }
            });
            std::copy(tmp.begin(), tmp.end(), prims.begin() + b);
            m = b + total_left;
        }
        if (m == b || m == e) return 0;
        return m;
    }

    // Median split on the widest centroid axis (or by index when centroids coincide).
    uint32_t median_split(uint32_t b, uint32_t e) {
        Box cen;
        bounds(b, e, &cen);
        int a = 0;
        float ext = -1;
        for (int k = 0; k < 3; ++k)
            if (cen.hi[k] - cen.lo[k] > ext) {
                ext = cen.hi[k] - cen.lo[k];
                a = k;
            }
        const uint32_t m = b + (e - b) / 2;
        if (ext > 0)
            std::nth_element(prims.data() + b, prims.data() + m, prims.data() + e,
                             [&](const BPrim& x, const BPrim& y) { return x.c[a] < y.c[a]; });
        return m;
    }

    uint32_t build(uint32_t b, uint32_t e, int depth) {
        const uint32_t me = next.fetch_add(1);
        if (me >= nodes.size()) throw std::logic_error("wide BVH node bound exceeded");
        struct Range {
            uint32_t b, e;
            bool leaf;
        };
        std::array<Range, 4> r{};
        int nr = 1;
        r[0] = {b, e, false};
        bool forced = depth >= kMaxWideDepth;
        while (nr < 4) {
            int pick = -1;
            uint32_t most = 0;
            for (int k = 0; k < nr; ++k)
                if (!r[k].leaf && r[k].e - r[k].b > most) {
                    most = r[k].e - r[k].b;
                    pick = k;
                }
            if (pick < 0) break;
            Range& g = r[pick];
            const uint32_t n = g.e - g.b;
            uint32_t m = 0;
            if (n > kLeafSize && !forced) m = sah_split(g.b, g.e);
            if (m == 0 && (n > kLeafMaxCount || (forced && n > kLeafSize))) m = median_split(g.b, g.e);
            if (m == 0) {
                g.leaf = true;
                continue;
            }
            const Range right{m, g.e, false};
            g.e = m;
            r[nr++] = right;
        }
        std::array<uint32_t, 4> child{kEmptyChild, kEmptyChild, kEmptyChild, kEmptyChild};
        std::array<Box, 4> boxes{};
        std::array<std::future<uint32_t>, 4> pending;
        for (int k = 0; k < nr; ++k) {
            boxes[k] = bounds(r[k].b, r[k].e, nullptr);
            const uint32_t n = r[k].e - r[k].b;
            const bool leaf = n <= kLeafSize || (r[k].leaf && n <= kLeafMaxCount);
            if (leaf) {
                if (r[k].b > kLeafMaxFirst) throw std::length_error("Too many triangles for the device BVH");
                child[k] = make_leaf(r[k].b, n);
                leaves.fetch_add(1);
            } else if (n > 65536) {
                const uint32_t cb = r[k].b, ce = r[k].e;
                pending[k] = std::async(std::launch::async, [this, cb, ce, depth] { return build(cb, ce, depth + 1); });
            } else {
                child[k] = build(r[k].b, r[k].e, depth + 1);
            }
        }
        for (int k = 0; k < nr; ++k)
            if (pending[k].valid()) child[k] = pending[k].get();
        WNode& node = nodes[me];
        for (int k = 0; k < 4; ++k) {
            const bool used = k < nr;
            node.lox[k] = used ? boxes[k].lo[0] : INFINITY;
            node.loy[k] = used ? boxes[k].lo[1] : INFINITY;
            node.loz[k] = used ? boxes[k].lo[2] : INFINITY;
            node.hix[k] = used ? boxes[k].hi[0] : -INFINITY;
            node.hiy[k] = used ? boxes[k].hi[1] : -INFINITY;
            node.hiz[k] = used ? boxes[k].hi[2] : -INFINITY;
            node.child[k] = child[k];
            node.pad[k] = 0;
        }
        return me;
    }
};

}  // namespace

// Early split clipping: a triangle whose box is far larger than the typical
// one (ground planes, walls spanning a scene) inflates every node above it.
// Its reference is split at the middle of the longest axis of its box, the
// triangle clipped to each half (in binary64), and each half's clipped box
// kept as a reference of its own, until a box is small enough (or 2^6
// references).  Boxes are widened by 2^-20 of their extent and rounded
// outward, so together they cover the triangle; the traversal tests the whole
// triangle in every reference it reaches (leaf-order primitives repeat the
// triangle; hit ordering is by triangle id, so repeats change no result).
namespace {

#ifndef CR_SPLIT_FACTOR
#define CR_SPLIT_FACTOR 4.0
#endif
constexpr double kSplitFactor = CR_SPLIT_FACTOR;  // split boxes above this multiple of the mean box area
#ifndef CR_SPLIT_DEPTH
#define CR_SPLIT_DEPTH 6
#endif
constexpr int kSplitDepth = CR_SPLIT_DEPTH;

using DPoint = std::array<double, 3>;

// The part of a convex polygon with coordinate `axis` <= c (keep_below) or >= c.
std::vector<DPoint> clip_half(const std::vector<DPoint>& poly, int axis, double c, bool keep_below) {
    std::vector<DPoint> out;
    const size_t k = poly.size();
    for (size_t i = 0; i < k; ++i) {
        const DPoint& a = poly[i];
        const DPoint& b = poly[(i + 1) % k];
        const double da = keep_below ? c - a[axis] : a[axis] - c, db = keep_below ? c - b[axis] : b[axis] - c;
        if (da >= 0) out.push_back(a);
        if ((da >= 0) != (db >= 0)) {
            const double t = da / (da - db);
            DPoint x;
            for (int j = 0; j < 3; ++j) x[j] = a[j] + (b[j] - a[j]) * t;
            x[axis] = c;
            out.push_back(x);
        }
    }
    return out;
}

void split_ref(const std::vector<DPoint>& poly, const Box& box, int depth, double limit, uint32_t id,
               std::vector<BPrim>& out) {
    if (depth >= kSplitDepth || box.area() <= limit) {
        BPrim p;
        for (int a = 0; a < 3; ++a) {
            p.lo[a] = box.lo[a];
            p.hi[a] = box.hi[a];
            p.c[a] = float(0.5 * (double(box.lo[a]) + double(box.hi[a])));
        }
        p.id = id;
        out.push_back(p);
        return;
    }
    int axis = 0;
    for (int a = 1; a < 3; ++a)
        if (double(box.hi[a]) - box.lo[a] > double(box.hi[axis]) - box.lo[axis]) axis = a;
    const double mid = 0.5 * (double(box.lo[axis]) + double(box.hi[axis]));
    for (int side = 0; side < 2; ++side) {
        const std::vector<DPoint> part = clip_half(poly, axis, mid, side == 0);
        if (part.size() < 3) continue;
        DPoint lo{INFINITY, INFINITY, INFINITY}, hi{-INFINITY, -INFINITY, -INFINITY};
        for (const DPoint& q : part)
            for (int a = 0; a < 3; ++a) {
                lo[a] = std::min(lo[a], q[a]);
                hi[a] = std::max(hi[a], q[a]);
            }
        Box sub;
        for (int a = 0; a < 3; ++a) {
            const double pad = (hi[a] - lo[a]) * 0x1p-20 + (std::fabs(lo[a]) + std::fabs(hi[a])) * 0x1p-40;
            sub.lo[a] = std::max(std::nextafter(float(lo[a] - pad), -INFINITY), box.lo[a]);
            sub.hi[a] = std::min(std::nextafter(float(hi[a] + pad), INFINITY), box.hi[a]);
        }
        split_ref(part, sub, depth + 1, limit, id, out);
    }
}

void split_large(const HostScene& s, std::vector<BPrim>& prims) {
    const size_t n = prims.size();
    double total = 0;
    for (const BPrim& p : prims) {
        Box b;
        b.grow(p.lo, p.hi);
        total += b.area();
    }
    const double limit = kSplitFactor * total / double(n);
    std::vector<BPrim> extra;
    for (size_t i = 0; i < n; ++i) {
        Box b;
        b.grow(prims[i].lo, prims[i].hi);
        if (!(b.area() > limit) || extra.size() > n) continue;
        const HTriangle& t = s.tris[prims[i].id];
        std::vector<DPoint> poly(3);
        for (int k = 0; k < 3; ++k) poly[k] = {t.p[k].x, t.p[k].y, t.p[k].z};
        std::vector<BPrim> refs;
        split_ref(poly, b, 0, limit, prims[i].id, refs);
        if (refs.empty()) continue;
        prims[i] = refs[0];
        extra.insert(extra.end(), refs.begin() + 1, refs.end());
    }
    prims.insert(prims.end(), extra.begin(), extra.end());
}

}  // namespace

namespace {

// ---- quantized device trees (layout.h QNode): per-axis origin and scale of a
// node, and child planes rounded outward in the decode arithmetic itself.
struct QAxis {
    float p;
    uint32_t e;
    float s;
};

// Smallest scale exponent whose 255 steps reach `hi` from `lo` when decoded
// (e = 254 decodes 255 steps to +inf, which covers any hi).
QAxis quant_axis(float lo, float hi) {
    QAxis a{lo, 1, qscale(1)};
    if (!(hi > lo)) return a;
    int e = 127 + int(std::ceil(std::log2((double(hi) - double(lo)) / 255.0)));
    e = std::clamp(e, 1, 254);
    while (e < 254 && !(qdecode(lo, 255, qscale(uint32_t(e))) >= hi)) ++e;
    while (e > 1 && qdecode(lo, 255, qscale(uint32_t(e - 1))) >= hi) --e;
    a.e = uint32_t(e);
    a.s = qscale(a.e);
    return a;
}

// Largest q with decode(q) <= v (q = 0 decodes to the origin, the node's low
// corner, which is <= v), and smallest q with decode(q) >= v.
uint32_t quant_lo(const QAxis& a, float v) {
    const double q = std::floor((double(v) - double(a.p)) / double(a.s));
    uint32_t k = q <= 0 ? 0u : (q >= 255 ? 255u : uint32_t(q));
    while (k > 0 && !(qdecode(a.p, k, a.s) <= v)) --k;
    return k;
}
uint32_t quant_hi(const QAxis& a, float v) {
    const double q = std::ceil((double(v) - double(a.p)) / double(a.s));
    uint32_t k = q <= 0 ? 0u : (q >= 255 ? 255u : uint32_t(q));
    while (k < 255 && !(qdecode(a.p, k, a.s) >= v)) ++k;
    return k;
}

struct QChild {
    uint32_t ref;
    Box box;
};

template <int W>
QNode<W> quantize_node(const QChild* kids, int n) {
    QNode<W> q{};
    Box all;
    for (int c = 0; c < n; ++c) all.grow(kids[c].box);
    QAxis ax[3];
    for (int a = 0; a < 3; ++a) ax[a] = n ? quant_axis(all.lo[a], all.hi[a]) : QAxis{0, 1, qscale(1)};
    q.px = ax[0].p;
    q.py = ax[1].p;
    q.pz = ax[2].p;
    q.exps = ax[0].e | (ax[1].e << 8) | (ax[2].e << 16) | (uint32_t(n) << 24);
    for (int c = 0; c < W; ++c) {
        q.child[c] = c < n ? kids[c].ref : kEmptyChild;
        for (int a = 0; a < 3; ++a) {
            const uint32_t lo = c < n ? quant_lo(ax[a], kids[c].box.lo[a]) : 0u;
            const uint32_t hi = c < n ? quant_hi(ax[a], kids[c].box.hi[a]) : 0u;
            q.qlo[a][c / 4] |= lo << (8 * (c % 4));
            q.qhi[a][c / 4] |= hi << (8 * (c % 4));
        }
    }
    return q;
}

int wnode_children(const WNode& w, QChild* out) {
    int n = 0;
    for (int c = 0; c < 4; ++c) {
        if (w.child[c] == kEmptyChild) continue;
        Box b;
        b.lo[0] = w.lox[c], b.lo[1] = w.loy[c], b.lo[2] = w.loz[c];
        b.hi[0] = w.hix[c], b.hi[1] = w.hiy[c], b.hi[2] = w.hiz[c];
        out[n++] = QChild{w.child[c], b};
    }
    return n;
}

// The 8-wide tree over the 4-wide tree's leaves: a node takes its 4-wide
// node's children, then repeatedly replaces the inner child of largest area
// by that child's own children while they fit in eight slots.  Parents are
// emitted before their children.
struct Collapse8 {
    const std::vector<WNode>& n4;
    std::vector<QNode8>& out;
    uint32_t emit(uint32_t i4) {
        QChild kids[8];
        int n = wnode_children(n4[i4], kids);
        for (;;) {
            int pick = -1;
            double best = -1;
            for (int c = 0; c < n; ++c) {
                if (ref_is_leaf(kids[c].ref)) continue;
                QChild sub[4];
                const int k = wnode_children(n4[kids[c].ref], sub);
                if (n - 1 + k > 8) continue;
                const double a = kids[c].box.area();
                if (a > best) {
                    best = a;
                    pick = c;
                }
            }
            if (pick < 0) break;
            QChild sub[4];
            const int k = wnode_children(n4[kids[pick].ref], sub);
            kids[pick] = sub[0];
            for (int j = 1; j < k; ++j) kids[n++] = sub[j];
        }
        const uint32_t me = uint32_t(out.size());
        out.emplace_back();
        for (int c = 0; c < n; ++c)
            if (!ref_is_leaf(kids[c].ref)) kids[c].ref = emit(kids[c].ref);
        out[me] = quantize_node<8>(kids, n);
        return me;
    }
};

}  // namespace

void unit_vector(const float* v, float* out) {
    const HVec3 u = unit_of(DV{v[0], v[1], v[2]});
    out[0] = u.x;
    out[1] = u.y;
    out[2] = u.z;
}

QNode4 quantize_boxes4(const float* lo, const float* hi, const uint32_t* refs, int n) {
    QChild kids[4];
    for (int c = 0; c < n && c < 4; ++c) {
        kids[c].ref = refs[c];
        for (int a = 0; a < 3; ++a) {
            kids[c].box.lo[a] = lo[3 * c + a];
            kids[c].box.hi[a] = hi[3 * c + a];
        }
    }
    return quantize_node<4>(kids, n < 4 ? n : 4);
}

// Node layout.  The parallel builder numbers nodes in the order threads
// create them, so siblings and parents and children lie far apart.  Relabel
// so that each node's inner children sit in consecutive slots (a 128-byte
// line holds two 64-byte quantized nodes), blocks in depth-first order from
// the root; every child still comes after its parent.
static void relabel_children_contiguous(std::vector<WNode>& nodes) {
    const size_t n = nodes.size();
    if (n < 2) return;
    std::vector<uint32_t> to(n, kEmptyChild);
    to[0] = 0;
    uint32_t next = 1;
    std::vector<uint32_t> stack{0};
    while (!stack.empty()) {
        const uint32_t i = stack.back();
        stack.pop_back();
        const WNode& w = nodes[i];
        for (int c = 0; c < 4; ++c)
            if (w.child[c] != kEmptyChild && !ref_is_leaf(w.child[c])) to[w.child[c]] = next++;
        for (int c = 3; c >= 0; --c)  // first child's block is laid out first
            if (w.child[c] != kEmptyChild && !ref_is_leaf(w.child[c])) stack.push_back(w.child[c]);
    }
    if (next != n) throw std::logic_error("wide BVH relabel: unreachable nodes");
    std::vector<WNode> out(n);
    for (size_t i = 0; i < n; ++i) {
        WNode w = nodes[i];
        for (int c = 0; c < 4; ++c)
            if (w.child[c] != kEmptyChild && !ref_is_leaf(w.child[c])) w.child[c] = to[w.child[c]];
        out[to[i]] = w;
    }
    nodes.swap(out);
}

void quantize_trees(WideBvh& b) {
    b.q4.resize(b.nodes.size());
    parallel_chunks(b.nodes.size(), 8192, [&](size_t lo, size_t hi, size_t) {
        for (size_t i = lo; i < hi; ++i) {
            QChild kids[4];
            const int n = wnode_children(b.nodes[i], kids);
            b.q4[i] = quantize_node<4>(kids, n);
        }
    });
    b.q8.clear();
    if (b.nodes.empty()) return;
    b.q8.reserve(b.nodes.size() / 2 + 1);
    Collapse8{b.nodes, b.q8}.emit(0);
}

WideBvh build_wide_bvh(const HostScene& s) {
    WideBvh out;
    const size_t n = s.tris.size();
    if (n == 0) return out;
    if (n > size_t(std::numeric_limits<uint32_t>::max() / 2))
        throw std::length_error("Too many triangles for 32-bit BVH node indices");
    std::vector<BPrim> prims(n);
    parallel_chunks(n, 65536, [&](size_t first, size_t last, size_t) {
    for (size_t i = first; i < last; ++i) {
        const HTriangle& t = s.tris[i];
        BPrim& p = prims[i];
        p.id = uint32_t(i);
        for (int a = 0; a < 3; ++a) {
            const float v0 = a == 0 ? t.p[0].x : (a == 1 ? t.p[0].y : t.p[0].z);
            const float v1 = a == 0 ? t.p[1].x : (a == 1 ? t.p[1].y : t.p[1].z);
            const float v2 = a == 0 ? t.p[2].x : (a == 1 ? t.p[2].y : t.p[2].z);
            p.lo[a] = std::min({v0, v1, v2});
            p.hi[a] = std::max({v0, v1, v2});
            p.c[a] = float((double(v0) + double(v1) + double(v2)) * (1.0 / 3.0));
        }
    }
    });
    split_large(s, prims);
    const size_t m = prims.size();
    out.nodes.resize(2 * m + 2);  // every wide node owns >= 1 leaf or >= 2 children
    Builder b{prims, out.nodes};
    b.build(0, uint32_t(m), 0);
    out.order.resize(m);
    for (size_t i = 0; i < m; ++i) out.order[i] = prims[i].id;
    out.nodes.resize(b.next.load());
    out.binary_nodes = 2 * b.leaves.load() - 1;
    relabel_children_contiguous(out.nodes);
    quantize_trees(out);
    return out;
}

// ------------------------------------------------------------------ environment
HostEnv build_env(const crg_env_desc& d) {
    HostEnv e;
    if (!d.rgba) {
        if (!fin3(d.constant) || std::min({d.constant[0], d.constant[1], d.constant[2]}) < 0)
            throw std::runtime_error("Invalid environment radiance");
        e.is_image = false;
        std::memcpy(e.constant, d.constant, sizeof(e.constant));
        e.integral = 4 * kHostPi * luminance3(d.constant);
        return e;
    }
    if (d.width <= 0 || d.height <= 0) throw std::runtime_error("Invalid environment image dimensions");
    if (!std::isfinite(d.rotation_degrees) || !std::isfinite(d.intensity) || d.intensity < 0)
        throw std::runtime_error("Invalid environment rotation/intensity");
    e.is_image = true;
    e.width = d.width;
    e.height = d.height;
    e.rotation = std::fmod(d.rotation_degrees, 360.0f) * kHostPi / 180;
    e.intensity = d.intensity;
    const size_t n = size_t(d.width) * size_t(d.height);
    e.texels.assign(d.rgba, d.rgba + 4 * n);
    std::vector<double> cdf;
    cdf.reserve(n);
    auto solid_angle = [&](int row) {
        const double a = double(kHostPi) * row / d.height, b = double(kHostPi) * (row + 1) / d.height;
        return (2 * double(kHostPi) / d.width) * (std::cos(a) - std::cos(b));
    };
    for (int y = 0; y < d.height; ++y) {
        const double omega = solid_angle(y);
        for (int x = 0; x < d.width; ++x) {
            const float* p = d.rgba + 4 * (size_t(y) * d.width + x);
            if (!fin3(p) || std::min({p[0], p[1], p[2]}) < 0)
                throw std::runtime_error("Non-finite or negative HDR environment pixel");
            const double light = luminance3(p) * e.intensity;
            e.integral += light * omega;
            e.weight_sum += (light + 1e-9) * omega;
            cdf.push_back(e.weight_sum);
        }
    }
    e.cdf = cdf;
    e.pdf.resize(n);
    for (size_t i = 0; i < n; ++i) {
        const int y = int(i / size_t(d.width));
        e.pdf[i] = float((cdf[i] - (i ? cdf[i - 1] : 0)) / (e.weight_sum * solid_angle(y)));
    }
    // T_i = first 24-bit draw k with pick(k) = double(k*2^-24)*W >= cdf[i]; the
    // texel picked for k is then #{i : T_i <= k} (upper_bound semantics).
    const double W = e.weight_sum;
    auto pick = [&](uint32_t k) { return double(float(k) * (1.0f / 16777216.0f)) * W; };
    e.threshold.resize(n);
    for (size_t i = 0; i < n; ++i) {
        uint32_t lo = 0, hi = 1u << 24;  // search in [0, 2^24]; 2^24 means "never"
        while (lo < hi) {
            const uint32_t mid = lo + (hi - lo) / 2;
            if (pick(mid) >= cdf[i]) hi = mid;
            else lo = mid + 1;
        }
        e.threshold[i] = lo;
    }
    // guide[j] = #{i : T_i <= j << kGuideShift} (thresholds are nondecreasing:
    // one merge pass), the last entry n.  It bounds the search window
    // [G[k >> s], G[(k >> s) + 1]] for any draw k.
    e.guide.resize(kGuideEntries);
    size_t below = 0;
    for (uint32_t j = 0; j + 1 < kGuideEntries; ++j) {
        const uint32_t k = j << kGuideShift;
        while (below < n && e.threshold[below] <= k) ++below;
        e.guide[j] = uint32_t(below);
    }
    e.guide[kGuideEntries - 1] = uint32_t(n);
    e.rowcos.resize(size_t(d.height) + 1);
    for (int y = 0; y <= d.height; ++y) e.rowcos[y] = std::cos(kHostPi * float(y) / float(d.height));
    return e;
}

// ------------------------------------------------------------------ lights
SceneLights scene_lights(const HostScene& s) {
    SceneLights sl;
    float lo[3] = {INFINITY, INFINITY, INFINITY}, hi[3] = {-INFINITY, -INFINITY, -INFINITY};
    for (const HTriangle& t : s.tris)
        for (int k = 0; k < 3; ++k) {
            const float v[3] = {t.p[k].x, t.p[k].y, t.p[k].z};
            for (int a = 0; a < 3; ++a) {
                lo[a] = std::min(lo[a], v[a]);
                hi[a] = std::max(hi[a], v[a]);
            }
        }
    const float ex = hi[0] - lo[0], ey = hi[1] - lo[1], ez = hi[2] - lo[2];
    sl.radius2 = s.tris.empty() ? 1 : std::max(1.0, double(ex * ex + ey * ey + ez * ez) * 0.25);
    for (const crg_light& l : s.lights) {
        sl.punctual_weight.push_back(double(luminance3(l.color)) * l.intensity *
                                     (l.type == 1 ? kHostPi * sl.radius2 : 4 * kHostPi));
        DevPunctual p{};
        p.type = uint32_t(l.type);
        std::memcpy(p.position, l.position, sizeof(p.position));
        std::memcpy(p.direction, l.direction, sizeof(p.direction));
        std::memcpy(p.color, l.color, sizeof(p.color));
        p.intensity = l.intensity;
        p.range = l.range;
        p.cos_outer = std::cos(l.outer_cone);
        p.cos_inner = std::cos(l.inner_cone);
        sl.punctual.push_back(p);
    }
    for (uint32_t i = 0; i < s.tris.size(); ++i) {
        const HTriangle& t = s.tris[i];
        const crg_material& m = s.materials[t.material];
        const double w = double(t.area) * kHostPi * luminance3(m.emission) * (m.double_sided ? 2 : 1);
        if (w > 0) {
            sl.emissive.push_back(i);
            sl.emissive_weight.push_back(w);
        }
    }
    return sl;
}

std::vector<uint32_t> light_ranks(const SceneLights& sl, const std::vector<uint32_t>& prim_to_tri, size_t triangles) {
    std::vector<uint32_t> tri_rank(triangles, kNoTriangle);
    uint32_t base = 0;
    for (double w : sl.punctual_weight)
        if (w > 0) ++base;
    for (size_t k = 0; k < sl.emissive.size(); ++k) tri_rank[sl.emissive[k]] = base + uint32_t(k);
    std::vector<uint32_t> rank(prim_to_tri.size());
    for (size_t p = 0; p < prim_to_tri.size(); ++p) rank[p] = tri_rank[prim_to_tri[p]];
    return rank;
}

LightTable build_light_table(const SceneLights& sl, const std::vector<uint32_t>& tri_to_prim, double env_integral) {
    struct Entry {
        uint32_t type, index;
        double cdf;
        float probability;
    };
    std::vector<Entry> entries;
    double total = 0;
    auto add = [&](uint32_t type, uint32_t index, double w) {
        if (w > 0) {
            total += w;
            entries.push_back({type, index, total, 0});
        }
    };
    for (uint32_t i = 0; i < sl.punctual_weight.size(); ++i) add(0, i, sl.punctual_weight[i]);
    for (size_t k = 0; k < sl.emissive.size(); ++k) add(1, tri_to_prim[sl.emissive[k]], sl.emissive_weight[k]);
    add(2, 0, env_integral * kHostPi * sl.radius2);
    LightTable lt;
    double last = 0, cumulative = 0;
    for (Entry& e : entries) {
        const double p = 0.95 * (e.cdf - last) / total + 0.05 / double(entries.size());
        last = e.cdf;
        cumulative += p;
        e.probability = float(p);
        e.cdf = cumulative;
        if (e.type == 2) lt.env_probability = e.probability;
    }
    if (!entries.empty()) entries.back().cdf = 1;
    for (const Entry& e : entries) {
        // lower_bound(cdf >= u) with u = k*2^-24 exactly: chosen iff k <= floor(cdf*2^24).
        const double scaled = std::floor(e.cdf * 16777216.0);
        const uint32_t th = scaled >= 16777216.0 ? 16777216u : (scaled < 0 ? 0u : uint32_t(scaled));
        lt.entries.push_back({e.type, e.index, th, e.probability});
    }
    return lt;
}

// ------------------------------------------------------------------ camera / options / materials
DevCamera prepare_camera(const crg_camera& c, int width, int height) {
    const HVec3 pos{c.position[0], c.position[1], c.position[2]}, fw{c.forward[0], c.forward[1], c.forward[2]},
        rt{c.right[0], c.right[1], c.right[2]}, up{c.up[0], c.up[1], c.up[2]};
    auto fin_v = [](HVec3 v) { return std::isfinite(v.x) && std::isfinite(v.y) && std::isfinite(v.z); };
    if (width <= 0 || height <= 0 || !fin_v(pos) || !fin_v(fw) || !fin_v(rt) || !fin_v(up) || !std::isfinite(c.yfov) ||
        c.yfov <= 0 || c.yfov >= kHostPi || !std::isfinite(c.aspect) || c.aspect < 0 || !std::isfinite(c.near_clip) ||
        c.near_clip < 0 || std::isnan(c.far_clip) || c.far_clip <= c.near_clip)
        throw std::invalid_argument("Invalid perspective camera or image dimensions");
    const HVec3 f = unit_of(dv(fw));
    const HVec3 r = orthogonal_of(dv(rt), f);
    const HVec3 u = orthogonal_of(dv(orthogonal_of(dv(up), f)), r);
    if (!nonzero3(f) || !nonzero3(r) || !nonzero3(u)) throw std::invalid_argument("Degenerate camera basis");
    DevCamera cam{};
    std::memcpy(cam.position, c.position, sizeof(cam.position));
    cam.f[0] = f.x; cam.f[1] = f.y; cam.f[2] = f.z;
    cam.r[0] = r.x; cam.r[1] = r.y; cam.r[2] = r.z;
    cam.u[0] = u.x; cam.u[1] = u.y; cam.u[2] = u.z;
    const double ratio = c.aspect > 0 ? double(c.aspect) : double(width) / height;
    const double half = std::tan(double(c.yfov) * 0.5);
    std::memcpy(&cam.ratio.b, &ratio, 8);
    std::memcpy(&cam.half_height.b, &half, 8);
    cam.near_clip = c.near_clip;
    cam.far_clip = c.far_clip;
    cam.width = width;
    cam.height = height;
    return cam;
}

std::string format_path_error(const ErrorRec& e, int width, int spp) {
    const uint64_t pixel = e.key / uint64_t(spp), sample = e.key % uint64_t(spp);
    const char* cause = "";
    switch (e.code) {
        case kErrNonFiniteThroughput: cause = "Non-finite path throughput"; break;
        case kErrBadRadiance: cause = "Non-finite or negative path radiance"; break;
        case kErrNested: cause = "Nested/intersecting refractive volumes are unsupported; separate the closed meshes"; break;
        case kErrInvalidSecondary: cause = "Invalid secondary ray"; break;
        case kErrOffsetOverflow: cause = "Secondary ray offset exceeds finite coordinates"; break;
        case kErrCamera: cause = "Unrepresentable perspective ray"; break;
        case kErrStack: cause = "BVH traversal exceeded its bounded stack"; break;
        case kErrInternal: cause = "Work item outside its render request"; break;
        case kErrTexture: cause = "Non-finite texture coordinates or transform"; break;
        default:
            cause = "Refractive exit without entry; check outward normals and keep the camera outside closed volumes";
    }
    std::ostringstream m;
    m << "Pixel (" << pixel % uint64_t(width) << "," << pixel / uint64_t(width) << "), sample " << sample << ": "
      << cause;
    if (e.code == kErrExitWithoutEntry) {
        double t;
        std::memcpy(&t, &e.tbits, 8);
        m.precision(9);
        m << "; bounce " << e.depth << ", triangle " << e.triangle << ", material " << e.material << ", t " << t
          << ", point (" << e.point[0] << ',' << e.point[1] << ',' << e.point[2] << "), origin (" << e.origin[0] << ','
          << e.origin[1] << ',' << e.origin[2] << "), direction (" << e.direction[0] << ',' << e.direction[1] << ','
          << e.direction[2] << ')';
    }
    return m.str();
}

void validate_options(const crg_options& o) {
    if (o.width < 1 || o.height < 1 || o.width > 16384 || o.height > 16384)
        throw std::runtime_error("Width and height must be in [1, 16384]");
    const uint64_t pixels = uint64_t(o.width) * uint64_t(o.height);
    if (pixels > 67108864) throw std::runtime_error("Image exceeds the 64-megapixel allocation limit");
    if (o.samples < 1 || o.samples > 1048576) throw std::runtime_error("Samples must be in [1, 1048576]");
    if (o.bounces < 1 || o.bounces > 64) throw std::runtime_error("Bounces must be in [1, 64]");
    if (o.threads < 0 || o.threads > 256) throw std::runtime_error("Threads must be 0 (automatic) or in [1, 256]");
}

std::vector<DevMaterial> device_materials(const HostScene& s) {
    std::vector<DevMaterial> out;
    for (const crg_material& m : s.materials) {
        DevMaterial d{};
        std::memcpy(d.base, m.base_color, sizeof(d.base));
        d.metallic = m.metallic;
        d.roughness = m.roughness;
        std::memcpy(d.emission, m.emission, sizeof(d.emission));
        d.transmission = m.transmission;
        d.ior = m.ior;
        d.thickness = m.thickness;
        d.att_distance = m.attenuation_distance;
        std::memcpy(d.att_color, m.attenuation_color, sizeof(d.att_color));
        if (!std::isfinite(m.attenuation_distance)) {
            d.sigma[0] = d.sigma[1] = d.sigma[2] = 0;
        } else {
            for (int k = 0; k < 3; ++k) {
                const float c = std::clamp(m.attenuation_color[k], 1e-8f, 1.0f);
                d.sigma[k] = -std::log(c) / m.attenuation_distance;
            }
        }
        d.double_sided = m.double_sided ? 1u : 0u;
        d.normal_scale = m.normal_scale;
        d.alpha_cutoff = m.alpha_cutoff;
        const crg_texture_slot* src[5] = {&m.base_color_tex, &m.metallic_roughness_tex, &m.normal_tex,
                                          &m.emissive_tex, &m.transmission_tex};
        DevSlot* dst[5] = {&d.base_tex, &d.mr_tex, &d.normal_tex, &d.emissive_tex, &d.transmission_tex};
        for (int k = 0; k < 5; ++k) {
            DevSlot& o = *dst[k];
            o.texture = src[k]->texture;
            o.texcoord = src[k]->texcoord;
            o.offset[0] = src[k]->offset[0];
            o.offset[1] = src[k]->offset[1];
            o.scale[0] = src[k]->scale[0];
            o.scale[1] = src[k]->scale[1];
            const double c = std::cos(double(src[k]->rotation)), sn = std::sin(double(src[k]->rotation));
            std::memcpy(&o.cos_bits, &c, 8);
            std::memcpy(&o.sin_bits, &sn, 8);
            if (o.texture >= 0) d.maps = 1;
        }
        out.push_back(d);
    }
    return out;
}

SceneArrays scene_arrays(const HostScene& hs) {
    SceneArrays a;
    a.bvh = build_wide_bvh(hs);
    const size_t n = a.bvh.order.size();  // leaf-order primitives: triangles and their split references
    a.tri_to_prim.assign(hs.tris.size(), kNoTriangle);
    for (size_t i = n; i-- > 0;) a.tri_to_prim[a.bvh.order[i]] = uint32_t(i);  // each triangle's first reference
    a.materials = device_materials(hs);
    bool any_map = false, any_normal_map = false;
    for (const crg_material& m : hs.materials) {
        for (const crg_texture_slot* sl :
             {&m.base_color_tex, &m.metallic_roughness_tex, &m.normal_tex, &m.emissive_tex, &m.transmission_tex})
            any_map = any_map || sl->texture >= 0;
        any_normal_map = any_normal_map || m.normal_tex.texture >= 0;
    }
    a.prims.resize(n);
    a.prim_ng.resize(n);
    a.aux.resize(n);
    if (any_map) a.uvs.resize(3 * n);
    if (any_normal_map) a.tangents.resize(3 * n);
    parallel_chunks(n, 65536, [&](size_t first, size_t last, size_t) {
    for (size_t i = first; i < last; ++i) {
        const HTriangle& t = hs.tris[a.bvh.order[i]];
        const crg_material& m = hs.materials[t.material];
        uint32_t flags = 0;
        if (!m.double_sided && m.transmission <= 0) flags |= kPrimCull;
        if (m.alpha_mode == 1) {
            if (m.base_color_tex.texture >= 0) flags |= kPrimMaskTexture;
            else if (!(m.base_color[3] * 1.0f >= m.alpha_cutoff)) flags |= kPrimNeverHit;
        }
        const uint32_t cls = m.transmission > 0 ? kShadeTransmissive : m.metallic > 0 ? kShadeMetal : kShadeOpaque;
        flags |= cls << kPrimClassShift;
        a.prims[i] = PrimRec{t.p[0].x, t.p[0].y, t.p[0].z, a.bvh.order[i], t.p[1].x, t.p[1].y, t.p[1].z, flags,
                             t.p[2].x, t.p[2].y, t.p[2].z, t.material};
        a.prim_ng[i] = make_float4(t.ng.x, t.ng.y, t.ng.z, 0);
        PrimAux& x = a.aux[i];
        const HVec3* nv[3] = {&t.n[0], &t.n[1], &t.n[2]};
        float* dst[3] = {x.n0, x.n1, x.n2};
        for (int k = 0; k < 3; ++k) {
            dst[k][0] = nv[k]->x;
            dst[k][1] = nv[k]->y;
            dst[k][2] = nv[k]->z;
            if (any_map) a.uvs[3 * i + k] = make_float4(t.uv[k][0], t.uv[k][1], t.uv[k][2], t.uv[k][3]);
            if (any_normal_map) a.tangents[3 * i + k] = make_float4(t.tan[k][0], t.tan[k][1], t.tan[k][2], t.tan[k][3]);
        }
        x.ng[0] = t.ng.x; x.ng[1] = t.ng.y; x.ng[2] = t.ng.z;
        x.ung[0] = t.ung.x; x.ung[1] = t.ung.y; x.ung[2] = t.ung.z;
        x.area = t.area;
    }
    });
    // Images sampled through a colour slot (base colour, emission) also get an
    // sRGB-decoded copy, decoded per texel with this host's powf exactly as the
    // reference decodes each fetched texel (reference src/math.h:61,
    // src/geometry.cpp:260-270), so the device never evaluates pow.
    std::vector<char> srgb_use(hs.images.size(), 0);
    for (const crg_material& m : hs.materials)
        for (const crg_texture_slot* slot : {&m.base_color_tex, &m.emissive_tex})
            if (slot->texture >= 0 && size_t(slot->texture) < hs.textures.size()) {
                const int32_t img = hs.textures[size_t(slot->texture)].image;
                if (img >= 0 && size_t(img) < srgb_use.size()) srgb_use[size_t(img)] = 1;
            }
    size_t total = 0;
    for (size_t i = 0; i < hs.images.size(); ++i) total += hs.images[i].texels.size() * (srgb_use[i] ? 2 : 1);
    a.texels.resize(total);
    // 8-bit texels (value k/255) take their decoded value from a table of the
    // 256 decodes; any other value is decoded directly.
    float srgb8[256];
    for (int k = 0; k < 256; ++k) srgb8[k] = srgb_to_linear(float(k) / 255.0f);
    const auto decode = [&](float v) {
        const float r = std::nearbyint(v * 255.0f);
        if (r >= 0.0f && r <= 255.0f && float(int(r)) / 255.0f == v && !std::signbit(v)) return srgb8[int(r)];
        return srgb_to_linear(v);
    };
    size_t at = 0;
    for (size_t i = 0; i < hs.images.size(); ++i) {
        const HImage& im = hs.images[i];
        DevImage d{at, at, im.width, im.height};
        float4* plain = a.texels.data() + at;
        at += im.texels.size();
        float4* dec = nullptr;
        if (srgb_use[i]) {
            d.srgb_offset = at;
            dec = a.texels.data() + at;
            at += im.texels.size();
        }
        parallel_chunks(im.texels.size(), 65536, [&](size_t b, size_t e, size_t) {
            std::memcpy(plain + b, im.texels.data() + b, (e - b) * sizeof(float4));
            if (dec)
                for (size_t k = b; k < e; ++k) {
                    const float4 v = im.texels[k];
                    dec[k] = make_float4(decode(v.x), decode(v.y), decode(v.z), v.w);
                }
        });
        a.images.push_back(d);
    }
    for (const crg_texture& t : hs.textures) a.textures.push_back(DevTexture{t.image, t.wrap_s, t.wrap_t, t.linear});
    return a;
}

}  // namespace cr
