// TEST INFRASTRUCTURE ONLY — the CPU oracle.
//
// A restatement of the reference renderer's hot path in plain C++ with native
// binary64, written from the reference's behaviour and kept in its floating-
// point expression order so results are bit-identical to the compiled
// reference.  Only tests/, __graft_entry__.smoke() and bench.py's cpu_baseline
// leg load it (through oracle/probe.py), never the product path.
//
// Follows (reference file:line):
//   Rng / sampleSeed / powerHeuristic ..... src/math.h:95-128
//   Bounds, intersectTriangle, offsetRay ... src/geometry.cpp:120-209
//   Camera::generate ........................ src/geometry.cpp:213-234
//   Scene::accepts / finalize ............... src/geometry.cpp:279-397
//   Scene::sample, interpolateUv ........... src/geometry.cpp:68-100, :238-275
//   surfaceAt ............................... src/geometry.cpp:401-459
//   Bvh::build / subdivide / intersect ...... src/geometry.cpp:463-639
//   BSDF .................................... src/bsdf.cpp:6-237
//   Environment, Lights, trace, render ...... src/render.cpp:12-456
//   writePng / writeHdr (film encode) ...... src/io/image.cpp:101-124, :201-236, with the
//                                            RGBE and run-length rules of the pinned
//                                            third_party/stb_image_write.h:635-790
#include <algorithm>
#include <array>
#include <atomic>
#include <chrono>
#include <cmath>
#include <cstring>
#include <limits>
#include <memory>
#include <mutex>
#include <numeric>
#include <sstream>
#include <stdexcept>
#include <string>
#include <thread>
#include <vector>

#include "probe.h"

namespace orc {

constexpr float PI = 3.14159265358979323846f;
constexpr float INF = std::numeric_limits<float>::infinity();
constexpr uint32_t NONE = std::numeric_limits<uint32_t>::max();

// ---------------------------------------------------------------- float vectors
struct V {
    float x = 0, y = 0, z = 0;
    float operator[](int i) const { return i == 0 ? x : (i == 1 ? y : z); }
    float& operator[](int i) { return i == 0 ? x : (i == 1 ? y : z); }
};
inline V vs(float s) { return {s, s, s}; }
inline V operator+(V a, V b) { return {a.x + b.x, a.y + b.y, a.z + b.z}; }
inline V operator-(V a, V b) { return {a.x - b.x, a.y - b.y, a.z - b.z}; }
inline V operator-(V a) { return {-a.x, -a.y, -a.z}; }
inline V operator*(V a, V b) { return {a.x * b.x, a.y * b.y, a.z * b.z}; }
inline V operator*(V a, float s) { return {a.x * s, a.y * s, a.z * s}; }
inline V operator*(float s, V a) { return a * s; }
inline V operator/(V a, float s) { return a * (1.0f / s); }
inline float dot(V a, V b) { return a.x * b.x + a.y * b.y + a.z * b.z; }
inline V cross(V a, V b) { return {a.y * b.z - a.z * b.y, a.z * b.x - a.x * b.z, a.x * b.y - a.y * b.x}; }
inline float len2(V a) { return dot(a, a); }
inline float len(V a) { return std::sqrt(len2(a)); }
inline V norm(V a) {
    const float l = len(a);
    return l > 0 ? a / l : V{};
}
inline bool fin(V a) { return std::isfinite(a.x) && std::isfinite(a.y) && std::isfinite(a.z); }
inline bool nz(V a) { return a.x != 0 || a.y != 0 || a.z != 0; }
inline float maxc(V a) { return std::max({a.x, a.y, a.z}); }
inline float lum(V a) { return dot(a, V{0.2126f, 0.7152f, 0.0722f}); }
inline V lerp(V a, V b, float t) { return a * (1 - t) + b * t; }
inline V refl(V i, V n) { return i - 2 * dot(i, n) * n; }
inline V vmin(V a, V b) { return {std::min(a.x, b.x), std::min(a.y, b.y), std::min(a.z, b.z)}; }
inline V vmax(V a, V b) { return {std::max(a.x, b.x), std::max(a.y, b.y), std::max(a.z, b.z)}; }
inline V vclamp(V a, float lo, float hi) {
    return {std::clamp(a.x, lo, hi), std::clamp(a.y, lo, hi), std::clamp(a.z, lo, hi)};
}
inline V v3(const float* p) { return {p[0], p[1], p[2]}; }
inline void put(float* o, V v) { o[0] = v.x; o[1] = v.y; o[2] = v.z; }

struct Basis {
    V x, y, z;
    explicit Basis(V n) : z(n) {
        x = norm(cross(std::abs(n.z) < 0.999f ? V{0, 0, 1} : V{0, 1, 0}, n));
        y = cross(n, x);
    }
    V local(V v) const { return {dot(v, x), dot(v, y), dot(v, z)}; }
    V world(V v) const { return x * v.x + y * v.y + z * v.z; }
};

// ---------------------------------------------------------------- double vectors
struct D {
    double x, y, z;
    double operator[](int i) const { return i == 0 ? x : (i == 1 ? y : z); }
};
inline D dd(V v) { return {v.x, v.y, v.z}; }
inline D operator+(D a, D b) { return {a.x + b.x, a.y + b.y, a.z + b.z}; }
inline D operator-(D a, D b) { return {a.x - b.x, a.y - b.y, a.z - b.z}; }
inline D operator*(D a, double s) { return {a.x * s, a.y * s, a.z * s}; }
inline double ddot(D a, D b) { return a.x * b.x + a.y * b.y + a.z * b.z; }
inline D dcross(D a, D b) { return {a.y * b.z - a.z * b.y, a.z * b.x - a.x * b.z, a.x * b.y - a.y * b.x}; }
inline V unitd(D v, V fb = {}) {
    const double s = ddot(v, v);
    if (!(s > 0) || !std::isfinite(s)) return fb;
    const double inv = 1 / std::sqrt(s);
    return {float(v.x * inv), float(v.y * inv), float(v.z * inv)};
}
inline V ortho(D v, V n, V fb = {}) {
    const double nn = ddot(dd(n), dd(n));
    if (!(nn > 0)) return fb;
    const D p = v - dd(n) * (ddot(v, dd(n)) / nn);
    const double tol = 4 * double(std::numeric_limits<float>::epsilon());
    if (ddot(p, p) <= ddot(v, v) * tol * tol) return fb;
    return unitd(p, fb);
}

// ---------------------------------------------------------------- random stream
struct Rng {
    uint64_t s;
    uint32_t next() {
        s += 0x9e3779b97f4a7c15ULL;
        uint64_t z = s;
        z = (z ^ (z >> 30)) * 0xbf58476d1ce4e5b9ULL;
        z = (z ^ (z >> 27)) * 0x94d049bb133111ebULL;
        return uint32_t((z ^ (z >> 31)) >> 32);
    }
    float uni() { return float(next() >> 8) * (1.0f / 16777216.0f); }
};
inline uint64_t seedOf(uint64_t seed, uint64_t pixel, uint64_t sample) {
    Rng h{seed ^ (pixel * 0xd2b74407b1ce6e93ULL) ^ (sample * 0xca5a826395121157ULL)};
    const uint64_t hi = h.next();
    return (hi << 32) | h.next();
}
inline float mis(float a, float b) {
    if (a <= 0) return 0;
    const double aa = double(a) * a, bb = double(b) * b;
    return float(aa / (aa + bb));
}

// ---------------------------------------------------------------- scene
struct Ray {
    V o, d;
    float tmin = 0, tmax = INF;
    V lo{};
};
struct Hit {
    double t = INF, u = 0, v = 0;
    uint32_t tri = 0;
};
struct Slot {
    int texture = -1, texcoord = 0;
    float offset[2] = {0, 0}, scale[2] = {1, 1}, rotation = 0;
};
struct Mat {
    V base{1, 1, 1};
    float alpha = 1, metallic = 1, roughness = 1;
    V emission{};
    float transmission = 0, ior = 1.5f, thickness = 0, attDist = INF;
    V attColor{1, 1, 1};
    bool twoSided = false, mask = false;
    float cutoff = 0.5f, normalScale = 1;
    Slot baseTex, mrTex, normalTex, emissiveTex, transmissionTex;
};
struct Image {
    int w = 0, h = 0;
    std::vector<std::array<float, 4>> px;
};
enum Wrap { REPEAT = 0, CLAMP = 1, MIRROR = 2 };
struct Texture {
    int image = -1, wrapS = REPEAT, wrapT = REPEAT;
    bool linear = true;
};
struct Tri {
    V p[3], n[3];
    float tan[3][4];
    float uv[3][2][2];
    uint32_t mat = 0;
    V ng;
    float area = 0;
};
struct Punctual {
    int type = 0;
    V pos, dir{0, 0, -1}, color{1, 1, 1};
    float intensity = 1, range = INF, inner = 0, outer = PI / 4;
};
struct Scene {
    std::vector<Tri> tris;
    std::vector<Mat> mats;
    std::vector<Punctual> lights;
    std::vector<Image> images;
    std::vector<Texture> textures;
};

// ---------------------------------------------------------------- textures
using F4 = std::array<float, 4>;
inline float srgbLinear(float v) { return v <= 0.04045f ? v / 12.92f : std::pow((v + 0.055f) / 1.055f, 2.4f); }

// UV reduction before texel indexing (reference src/geometry.cpp:77-84).
double wrapCoord(double x, int wrap) {
    if (wrap == CLAMP) return std::clamp(x, 0.0, 1.0);
    if (wrap == REPEAT) return x - std::floor(x);
    double r = std::fmod(x, 2.0);
    if (r < 0) r += 2;
    return r <= 1 ? r : 2 - r;
}
// Integer neighbour wrapping (reference src/geometry.cpp:87-95).
int wrapIndex(int64_t i, int size, int wrap) {
    const int64_t n = size;
    if (wrap == CLAMP) return int(std::clamp(i, int64_t{0}, n - 1));
    const int64_t period = wrap == MIRROR ? 2 * n : n;
    i %= period;
    if (i < 0) i += period;
    if (wrap == MIRROR && i >= n) i = period - 1 - i;
    return int(i);
}
F4 lerp4(F4 a, F4 b, float t) {
    return {a[0] * (1 - t) + b[0] * t, a[1] * (1 - t) + b[1] * t, a[2] * (1 - t) + b[2] * t, a[3] * (1 - t) + b[3] * t};
}
// Scene::sample (reference src/geometry.cpp:238-275): scale, rotate, translate,
// wrap, then nearest or bilinear with per-texel sRGB decode.
F4 sampleTex(const Scene& sc, const Slot& slot, const float uv0[2], const float uv1[2], bool srgb) {
    if (slot.texture == -1) return {1, 1, 1, 1};
    const Texture& tex = sc.textures[size_t(slot.texture)];
    const Image& im = sc.images[size_t(tex.image)];
    const float* uv = slot.texcoord == 0 ? uv0 : uv1;
    if (!std::isfinite(uv[0]) || !std::isfinite(uv[1]) || !std::isfinite(slot.offset[0]) ||
        !std::isfinite(slot.offset[1]) || !std::isfinite(slot.scale[0]) || !std::isfinite(slot.scale[1]) ||
        !std::isfinite(slot.rotation))
        throw std::runtime_error("Non-finite texture coordinates or transform");
    const double c = std::cos(double(slot.rotation)), s = std::sin(double(slot.rotation));
    const double x = double(uv[0]) * slot.scale[0], y = double(uv[1]) * slot.scale[1];
    const double u = wrapCoord(c * x - s * y + slot.offset[0], tex.wrapS);
    const double v = wrapCoord(s * x + c * y + slot.offset[1], tex.wrapT);
    auto texel = [&](int64_t ix, int64_t iy) {
        const int tx = wrapIndex(ix, im.w, tex.wrapS), ty = wrapIndex(iy, im.h, tex.wrapT);
        F4 val = im.px[size_t(ty) * size_t(im.w) + size_t(tx)];
        if (srgb)
            for (int k = 0; k < 3; ++k) val[k] = srgbLinear(val[k]);
        return val;
    };
    if (!tex.linear) return texel(int64_t(std::floor(u * im.w)), int64_t(std::floor(v * im.h)));
    const double px = u * im.w - 0.5, py = v * im.h - 0.5;
    const int64_t ix = int64_t(std::floor(px)), iy = int64_t(std::floor(py));
    const float fx = float(px - std::floor(px)), fy = float(py - std::floor(py));
    return lerp4(lerp4(texel(ix, iy), texel(ix + 1, iy), fx), lerp4(texel(ix, iy + 1), texel(ix + 1, iy + 1), fx), fy);
}

bool validRay(const Ray& r) {
    return fin(r.o) && fin(r.lo) && fin(r.d) && nz(r.d) && !std::isnan(r.tmin) && !std::isnan(r.tmax) &&
           r.tmin <= r.tmax;
}

// Watertight shear test in binary64 (reference src/geometry.cpp:147-182).
bool hitTriangle(const Tri& t, const Ray& r, Hit& h) {
    if (!validRay(r) || std::isnan(h.t)) return false;
    const double lim = std::min(double(r.tmax), h.t);
    if (lim < r.tmin) return false;
    int kz = 0;
    if (std::abs(r.d.y) > std::abs(r.d[kz])) kz = 1;
    if (std::abs(r.d.z) > std::abs(r.d[kz])) kz = 2;
    int kx = (kz + 1) % 3, ky = (kx + 1) % 3;
    if (r.d[kz] < 0) std::swap(kx, ky);
    const double sx = double(r.d[kx]) / r.d[kz], sy = double(r.d[ky]) / r.d[kz], sz = 1.0 / r.d[kz];
    const D o = dd(r.o) + dd(r.lo);
    const D a = dd(t.p[0]) - o, b = dd(t.p[1]) - o, c = dd(t.p[2]) - o;
    const double ax = a[kx] - sx * a[kz], ay = a[ky] - sy * a[kz];
    const double bx = b[kx] - sx * b[kz], by = b[ky] - sy * b[kz];
    const double cx = c[kx] - sx * c[kz], cy = c[ky] - sy * c[kz];
    const double w0 = cx * by - cy * bx, w1 = ax * cy - ay * cx, w2 = bx * ay - by * ax;
    const bool anyNeg = w0 < 0 || w1 < 0 || w2 < 0, anyPos = w0 > 0 || w1 > 0 || w2 > 0;
    if (anyNeg && anyPos) return false;
    const double det = w0 + w1 + w2;
    if (det == 0 || !std::isfinite(det)) return false;
    const double t_ = (w0 * (sz * a[kz]) + w1 * (sz * b[kz]) + w2 * (sz * c[kz])) / det;
    if (!std::isfinite(t_) || std::abs(t_) > std::numeric_limits<float>::max() || t_ < r.tmin || t_ > lim)
        return false;
    h.t = t_;
    h.u = w1 / det;
    h.v = w2 / det;
    return true;
}

// interpolateUv (reference src/geometry.cpp:68-74)
void interpUv(const Tri& t, const Hit& h, int set, float out[2]) {
    const double w = 1.0 - double(h.u) - h.v;
    for (int k = 0; k < 2; ++k)
        out[k] = float(w * t.uv[0][set][k] + double(h.u) * t.uv[1][set][k] + double(h.v) * t.uv[2][set][k]);
}

bool accepted(const Scene& s, const Tri& t, const Ray& r, const Hit& h) {
    const Mat& m = s.mats[t.mat];
    if (!m.twoSided && m.transmission <= 0) {
        V n = t.ng;
        if (!nz(n)) n = unitd(dcross(dd(t.p[1]) - dd(t.p[0]), dd(t.p[2]) - dd(t.p[0])));
        if (ddot(dd(n), dd(r.d)) >= 0) return false;
    }
    if (m.mask) {
        float uv0[2], uv1[2];
        interpUv(t, h, 0, uv0);
        interpUv(t, h, 1, uv1);
        const F4 c = sampleTex(s, m.baseTex, uv0, uv1, false);
        return m.alpha * c[3] >= m.cutoff;
    }
    return true;
}

Ray offset(V p, V ng, V dir, float maxd) {
    if (!fin(p) || !fin(ng) || !fin(dir) || !nz(ng) || !nz(dir) || std::isnan(maxd) || maxd < 0)
        throw std::invalid_argument("Invalid secondary ray");
    dir = unitd(dd(dir));
    V n = unitd(dd(ng));
    if (ddot(dd(n), dd(dir)) < 0) n = -n;
    const double scale = std::max({std::abs(double(p.x)), std::abs(double(p.y)), std::abs(double(p.z))});
    const double dist = 4 * std::numeric_limits<float>::epsilon() * scale;
    V o = p;
    for (int a = 0; a < 3; ++a) {
        o[a] = float(double(p[a]) + dist * n[a]);
        if (n[a] != 0) o[a] = std::nextafter(o[a], n[a] > 0 ? INF : -INF);
    }
    if (!fin(o)) throw std::overflow_error("Secondary ray offset exceeds finite coordinates");
    float lim = maxd;
    if (std::isfinite(lim)) {
        const double adv = ddot(dd(o) - dd(p), dd(dir));
        lim = float(std::max(0.0, double(lim) - adv));
        lim = std::nextafter(lim, 0.0f);
    }
    return {o, dir, 0, lim, {}};
}

struct Cam {
    V pos, fwd{0, 0, -1}, right{1, 0, 0}, up{0, 1, 0};
    float yfov = PI / 4, aspect = 0, nearc = 0.01f, farc = INF;
};

Ray camRay(const Cam& c, float px, float py, int w, int h) {
    if (w <= 0 || h <= 0 || !std::isfinite(px) || !std::isfinite(py) || !fin(c.pos) || !fin(c.fwd) ||
        !fin(c.right) || !fin(c.up) || !std::isfinite(c.yfov) || c.yfov <= 0 || c.yfov >= PI ||
        !std::isfinite(c.aspect) || c.aspect < 0 || !std::isfinite(c.nearc) || c.nearc < 0 || std::isnan(c.farc) ||
        c.farc <= c.nearc)
        throw std::invalid_argument("Invalid perspective camera or image dimensions");
    const V f = unitd(dd(c.fwd));
    const V r = ortho(dd(c.right), f);
    const V u = ortho(dd(ortho(dd(c.up), f)), r);
    if (!nz(f) || !nz(r) || !nz(u)) throw std::invalid_argument("Degenerate camera basis");
    const double ratio = c.aspect > 0 ? c.aspect : double(w) / h;
    const double hh = std::tan(double(c.yfov) * 0.5);
    const double x = (2 * double(px) / w - 1) * ratio * hh;
    const double y = (1 - 2 * double(py) / h) * hh;
    const V d = unitd(dd(f) + dd(r) * x + dd(u) * y);
    const double proj = ddot(dd(d), dd(f));
    const double nt = c.nearc / proj;
    if (!(proj > 0) || !std::isfinite(nt) || nt > std::numeric_limits<float>::max())
        throw std::invalid_argument("Unrepresentable perspective ray");
    return {c.pos, d, float(nt), float(c.farc / proj), {}};
}

// ---------------------------------------------------------------- finalize
void finalize(Scene& s) {
    for (size_t i = 0; i < s.mats.size(); ++i) {
        const Mat& m = s.mats[i];
        auto unit01 = [](float v) { return std::isfinite(v) && v >= 0 && v <= 1; };
        if (!unit01(m.base.x) || !unit01(m.base.y) || !unit01(m.base.z) || !unit01(m.alpha) || !unit01(m.metallic) ||
            !unit01(m.roughness) || !unit01(m.transmission) || !unit01(m.cutoff) || !fin(m.emission) ||
            m.emission.x < 0 || m.emission.y < 0 || m.emission.z < 0 || !std::isfinite(m.ior) || m.ior <= 0 ||
            !std::isfinite(m.thickness) || m.thickness < 0 || std::isnan(m.attDist) || m.attDist <= 0 ||
            !unit01(m.attColor.x) || !unit01(m.attColor.y) || !unit01(m.attColor.z))
            throw std::runtime_error("Material " + std::to_string(i) + " has non-finite or out-of-range parameters");
    }
    size_t kept = 0;
    for (size_t i = 0; i < s.tris.size(); ++i) {
        Tri t = s.tris[i];
        if (t.mat >= s.mats.size())
            throw std::runtime_error("Triangle " + std::to_string(i) + " has an invalid material index");
        const D cr = dcross(dd(t.p[1]) - dd(t.p[0]), dd(t.p[2]) - dd(t.p[0]));
        const double twice = std::sqrt(ddot(cr, cr));
        if (twice * 0.5 > std::numeric_limits<float>::max())
            throw std::runtime_error("Triangle " + std::to_string(i) + " area exceeds the supported finite range");
        t.area = float(twice * 0.5);
        if (!(t.area > 0)) continue;
        t.ng = unitd(cr);
        // UV-derived tangent directions in the normal map's UV set (reference src/geometry.cpp:98-116)
        const Mat& tm = s.mats[t.mat];
        const int set = tm.normalTex.texture >= 0 ? tm.normalTex.texcoord : 0;
        D ut{0, 0, 0}, ub{0, 0, 0};
        bool usable = false;
        {
            const double du1 = double(t.uv[1][set][0]) - t.uv[0][set][0], dv1 = double(t.uv[1][set][1]) - t.uv[0][set][1];
            const double du2 = double(t.uv[2][set][0]) - t.uv[0][set][0], dv2 = double(t.uv[2][set][1]) - t.uv[0][set][1];
            const double det = du1 * dv2 - dv1 * du2;
            if (det != 0) {
                const D e1 = dd(t.p[1]) - dd(t.p[0]), e2 = dd(t.p[2]) - dd(t.p[0]);
                const double sg = det < 0 ? -1 : 1;
                ut = (e1 * dv2 - e2 * dv1) * sg;
                ub = (e2 * du1 - e1 * du2) * sg;
                usable = nz(unitd(ut)) && nz(unitd(ub));
            }
        }
        for (int k = 0; k < 3; ++k) {
            V n = unitd(dd(t.n[k]));
            float hand = t.tan[k][3] < 0 ? -1.0f : 1.0f;
            if (!nz(n) || dot(n, t.ng) == 0) n = t.ng;
            else if (dot(n, t.ng) < 0) {
                n = -n;
                hand = -hand;
            }
            V tg = ortho(D{t.tan[k][0], t.tan[k][1], t.tan[k][2]}, n);
            if (!nz(tg)) {
                if (usable) tg = ortho(ut, n);
                if (nz(tg)) hand = ddot(dd(cross(n, tg)), ub) < 0 ? -1.0f : 1.0f;
                else {
                    tg = Basis(n).x;
                    hand = 1;
                }
            }
            t.n[k] = n;
            t.tan[k][0] = tg.x;
            t.tan[k][1] = tg.y;
            t.tan[k][2] = tg.z;
            t.tan[k][3] = hand;
        }
        s.tris[kept++] = t;
    }
    s.tris.resize(kept);
}

// ---------------------------------------------------------------- surface
struct Surf {
    V p, plo, ng, n, tan, bit, base, emission, attColor;
    float metallic = 0, roughness = 1, transmission = 0, ior = 1.5f, thickness = 0, attDist = INF;
    uint32_t mat = 0;
    bool front = true, twoSided = false;
};

Surf surfaceAt(const Scene& sc, const Ray& r, const Hit& h) {
    if (h.tri >= sc.tris.size() || !std::isfinite(h.t) || !std::isfinite(h.u) || !std::isfinite(h.v))
        throw std::out_of_range("Invalid surface hit");
    const Tri& t = sc.tris[h.tri];
    const Mat& m = sc.mats[t.mat];
    const double w = 1.0 - h.u - h.v;
    const D P = dd(t.p[0]) * w + dd(t.p[1]) * h.u + dd(t.p[2]) * h.v;
    Surf s;
    s.p = {float(P.x), float(P.y), float(P.z)};
    s.plo = {float(P.x - s.p.x), float(P.y - s.p.y), float(P.z - s.p.z)};
    s.ng = t.ng;
    s.front = ddot(dd(r.d), dd(s.ng)) < 0;
    s.n = unitd(dd(t.n[0]) * w + dd(t.n[1]) * h.u + dd(t.n[2]) * h.v, s.ng);
    if (dot(s.n, s.ng) < 0) s.n = -s.n;
    auto tv = [&](int k) { return D{t.tan[k][0], t.tan[k][1], t.tan[k][2]}; };
    s.tan = ortho(tv(0) * w + tv(1) * h.u + tv(2) * h.v, s.n, Basis(s.n).x);
    const float hand = w * t.tan[0][3] + h.u * t.tan[1][3] + h.v * t.tan[2][3] < 0 ? -1.0f : 1.0f;
    float hand2 = hand;
    s.bit = cross(s.n, s.tan) * hand2;
    float uv0[2], uv1[2];
    interpUv(t, h, 0, uv0);
    interpUv(t, h, 1, uv1);
    const F4 base = sampleTex(sc, m.baseTex, uv0, uv1, true);
    s.base = m.base * V{base[0], base[1], base[2]};
    const F4 mr = sampleTex(sc, m.mrTex, uv0, uv1, false);
    s.metallic = std::clamp(m.metallic * mr[2], 0.0f, 1.0f);
    s.roughness = std::clamp(m.roughness * mr[1], 0.0f, 1.0f);
    const F4 em = sampleTex(sc, m.emissiveTex, uv0, uv1, true);
    s.emission = m.emission * V{em[0], em[1], em[2]};
    s.transmission = std::clamp(m.transmission * sampleTex(sc, m.transmissionTex, uv0, uv1, false)[0], 0.0f, 1.0f);
    s.ior = m.ior;
    s.thickness = m.thickness;
    s.attDist = m.attDist;
    s.attColor = m.attColor;
    s.mat = t.mat;
    s.twoSided = m.twoSided;
    if (m.normalTex.texture >= 0) {  // normal mapping (reference src/geometry.cpp:436-457)
        const Slot& sl = m.normalTex;
        if (sl.scale[0] != 0 && sl.scale[1] != 0) {
            const double c = std::cos(double(sl.rotation)), sn = std::sin(double(sl.rotation));
            const D tt = dd(s.tan) * (c / sl.scale[0]) - dd(s.bit) * (sn / sl.scale[1]);
            s.tan = ortho(tt, s.n, s.tan);
            if ((sl.scale[0] < 0) != (sl.scale[1] < 0)) hand2 = -hand2;
            s.bit = cross(s.n, s.tan) * hand2;
        }
        const F4 tx = sampleTex(sc, sl, uv0, uv1, false);
        const D mapped = dd(s.tan) * ((2.0 * tx[0] - 1) * m.normalScale) + dd(s.bit) * ((2.0 * tx[1] - 1) * m.normalScale) +
                         dd(s.n) * (2.0 * tx[2] - 1);
        s.n = unitd(mapped, s.n);
        if (dot(s.n, s.ng) < 0) s.n = -s.n;
        s.tan = ortho(dd(s.tan), s.n, Basis(s.n).x);
        s.bit = cross(s.n, s.tan) * hand2;
    }
    return s;
}

// ---------------------------------------------------------------- BSDF
float fresnel(float c, float ei, float et) {
    if (ei == et) return 0;
    c = std::clamp(std::abs(c), 0.0f, 1.0f);
    const float eta = ei / et;
    const float s2 = eta * eta * std::max(0.0f, 1 - c * c);
    if (s2 >= 1) return 1;
    const float ct = std::sqrt(1 - s2);
    const float ds = ei * c + et * ct, dp = et * c + ei * ct;
    if (ds == 0 || dp == 0) return 1;
    const float rs = (ei * c - et * ct) / ds, rp = (et * c - ei * ct) / dp;
    return 0.5f * (rs * rs + rp * rp);
}
bool refract(V wo, V n, float ratio, V& wi) {
    const float ci = dot(wo, n);
    const float s2 = ratio * ratio * std::max(0.0f, 1 - ci * ci);
    if (s2 >= 1) return false;
    wi = norm(-wo * ratio + n * (ratio * ci - std::sqrt(1 - s2)));
    return true;
}
V absorb(V c, float d) {
    if (!std::isfinite(d)) return {};
    if (!(d > 0)) throw std::runtime_error("Attenuation distance must be positive");
    c = vclamp(c, 1e-8f, 1);
    return {-std::log(c.x) / d, -std::log(c.y) / d, -std::log(c.z) / d};
}
V beer(V k, float d) {
    if (d == INF) return {k.x > 0 ? 0.0f : 1.0f, k.y > 0 ? 0.0f : 1.0f, k.z > 0 ? 0.0f : 1.0f};
    const V a = k * (-std::max(0.0f, d));
    return {std::exp(a.x), std::exp(a.y), std::exp(a.z)};
}
float ggxD(float al, V h) {
    if (h.z <= 0) return 0;
    const float a2 = al * al;
    const float q = (h.x * h.x + h.y * h.y) + a2 * h.z * h.z;
    return a2 / (PI * q * q);
}
float g1(float al, V v) {
    const float c2 = v.z * v.z;
    if (c2 == 0) return 0;
    return 2 / (1 + std::sqrt(1 + al * al * (v.x * v.x + v.y * v.y) / c2));
}
V vndf(V wo, float al, float u, float v) {
    const V vh = norm({al * wo.x, al * wo.y, wo.z});
    const float l2 = vh.x * vh.x + vh.y * vh.y;
    const V t1 = l2 > 0 ? V{-vh.y, vh.x, 0} / std::sqrt(l2) : V{1, 0, 0};
    const V t2 = cross(vh, t1);
    const float r = std::sqrt(u), phi = 2 * PI * v;
    const float x = r * std::cos(phi);
    float y = r * std::sin(phi);
    const float s = 0.5f * (1 + vh.z);
    y = (1 - s) * std::sqrt(std::max(0.0f, 1 - x * x)) + s * y;
    const V nh = t1 * x + t2 * y + vh * std::sqrt(std::max(0.0f, 1 - x * x - y * y));
    return norm({al * nh.x, al * nh.y, std::max(0.0f, nh.z)});
}

struct Lobes {
    const Surf& s;
    Basis fr;
    V wo;
    float ei, et, eta, al, pd, pdt = 0;
    bool thin, smooth, matched = false;
    Lobes(const Surf& sf, V out)
        : s(sf), fr(dot(out, sf.n) > 0 ? sf.n : -sf.n), wo(fr.local(out)),
          ei(sf.thickness > 0 && !sf.front ? sf.ior : 1), et(sf.thickness > 0 && !sf.front ? 1 : sf.ior),
          eta(et / ei), al(std::max(0.001f, sf.roughness * sf.roughness)), pd(0),
          thin(sf.thickness <= 0 && sf.transmission > 0), smooth(sf.roughness <= 0.001f) {
        if (thin) {
            ei = 1;
            et = sf.ior;
            eta = sf.ior;
        }
        matched = std::abs(ei - et) <= std::sqrt(std::numeric_limits<float>::epsilon()) * std::max(ei, et);
        if (matched) ei = et = eta = 1;
        const float f = F(wo.z);
        const float dif = (1 - s.metallic) * (1 - s.transmission) * lum(s.base) * (1 - f);
        const float tr = (1 - s.metallic) * s.transmission * (1 - f) * lum(s.base);
        const float tot = dif + lum(R(f, wo.z)) + tr;
        if (tot > 0) {
            pd = dif / tot;
            if (matched) pdt = tr / tot;
        }
    }
    float F(float c) const {
        const float f = fresnel(c, ei, et);
        return thin ? 2 * f / (1 + f) : f;
    }
    V R(float f, float c) const {
        const float sch = std::pow(1 - std::clamp(std::abs(c), 0.0f, 1.0f), 5.0f);
        return vs((1 - s.metallic) * f) + (s.base + (vs(1) - s.base) * sch) * s.metallic;
    }
    V T(float f) const { return s.base * ((1 - s.metallic) * s.transmission * (1 - f)); }
    float pr(float f, float ch) const {
        if (matched) return 1;
        const float r = lum(R(f, ch)), t = lum(T(f));
        return r + t > 0 ? r / (r + t) : 1;
    }
    float pm() const { return std::max(0.0f, 1 - pd - pdt); }
};

struct Eval {
    V value{};
    float pdf = 0;
};

Eval evalL(const Lobes& c, V wi) {
    Eval e;
    const float no = c.wo.z, ni = std::abs(wi.z);
    if (no <= 1e-7f || ni <= 1e-7f) return e;
    const bool up = wi.z > 0;
    if (!up && (c.matched || c.thin)) return e;
    if (up) {
        const float dif = (1 - c.s.metallic) * (1 - c.s.transmission) * (1 - c.F(no)) * (1 - c.F(ni));
        e.value = c.s.base * (dif / PI);
        e.pdf = c.pd * ni / PI;
    }
    if (c.smooth) return e;
    V h = norm(up ? c.wo + wi : c.wo + wi * c.eta);
    if (h.z < 0) h = -h;
    const float oh = dot(c.wo, h), ih = dot(wi, h);
    if (oh <= 0 || (up ? ih <= 0 : ih >= 0)) return e;
    const float f = c.F(oh), p = c.pr(f, oh);
    const float d = ggxD(c.al, h), ga = g1(c.al, c.wo), g = ga * g1(c.al, wi);
    const float ph = d * ga * oh / no;
    if (up) {
        e.value = e.value + c.R(f, oh) * (d * g / (4 * no * ni));
        e.pdf += c.pm() * p * ph / (4 * oh);
    } else {
        const float den = oh + c.eta * ih, den2 = den * den;
        if (den2 < 1e-20f) return {};
        e.value = c.T(f) * (d * g * std::abs(oh * ih) / (no * ni * den2));
        e.pdf = c.pm() * (1 - p) * ph * c.eta * c.eta * std::abs(ih) / den2;
    }
    return e;
}
bool side(const Surf& s, V wo, V wi, bool tr) {
    const float d = dot(wo, s.ng) * dot(wi, s.ng);
    return tr ? d < 0 : d > 0;
}
Eval evalBsdf(const Surf& s, V wo, V wi) {
    Lobes c(s, wo);
    const V l = c.fr.local(wi);
    if (!side(s, wo, wi, l.z < 0)) return {};
    return evalL(c, l);
}
struct Smp {
    V dir{}, w{};
    float pdf = 0, eta = 1;
    bool delta = false, tr = false;
    bool ok() const { return maxc(w) > 0 && fin(w) && fin(dir); }
};
Smp sampleBsdf(const Surf& s, V wo, Rng& g) {
    Lobes c(s, wo);
    Smp out;
    if (c.wo.z <= 1e-7f) return out;
    V wi;
    const float lobe = g.uni();
    if (lobe < c.pd) {
        const float u = g.uni(), phi = 2 * PI * g.uni();
        wi = {std::sqrt(u) * std::cos(phi), std::sqrt(u) * std::sin(phi), std::sqrt(1 - u)};
    } else if (lobe < c.pd + c.pdt) {
        out.dir = -wo;
        out.delta = out.tr = true;
        out.pdf = c.pdt;
        out.w = s.base * ((1 - s.metallic) * s.transmission / out.pdf);
        return side(s, wo, out.dir, true) ? out : Smp{};
    } else {
        const float u = g.uni(), v = g.uni();
        const V h = c.smooth ? V{0, 0, 1} : vndf(c.wo, c.al, u, v);
        const float oh = dot(c.wo, h), f = c.F(oh), p = c.pr(f, oh);
        const bool up = g.uni() < p;
        if (up) wi = refl(-c.wo, h);
        else if (c.thin) wi = -c.wo;
        else if (!refract(c.wo, h, 1 / c.eta, wi)) return out;
        if ((up && wi.z <= 0) || (!up && wi.z >= 0)) return out;
        out.tr = !up;
        out.eta = out.tr && !c.thin ? c.eta : 1;
        if (c.smooth || (!up && c.thin)) {
            out.delta = true;
            out.dir = norm(c.fr.world(wi));
            if (!side(s, wo, out.dir, out.tr)) return {};
            const float pb = c.pm() * (up ? p : 1 - p);
            if (pb <= 0) return {};
            out.pdf = pb;
            out.w = (up ? c.R(f, oh) : c.T(f) / (out.eta * out.eta)) / pb;
            return out;
        }
    }
    out.dir = norm(c.fr.world(wi));
    wi = c.fr.local(out.dir);
    if ((wi.z < 0) != out.tr || !side(s, wo, out.dir, out.tr)) return {};
    const Eval e = evalL(c, wi);
    if (e.pdf <= 0 || !std::isfinite(e.pdf)) return {};
    out.pdf = e.pdf;
    out.w = e.value * (std::abs(wi.z) / e.pdf);
    return out;
}

// ---------------------------------------------------------------- environment
struct Env {
    V constant{};
    int w = 0, h = 0;
    std::vector<V> px;
    float rot = 0, inten = 1;
    double integral = 0, wsum = 0;
    std::vector<double> cdf;

    double omega(int row) const {
        const double a = double(PI) * row / h, b = double(PI) * (row + 1) / h;
        return (2 * double(PI) / w) * (std::cos(a) - std::cos(b));
    }
    void init(const crg_env_desc& d) {
        if (!d.rgba) {
            constant = v3(d.constant);
            integral = 4 * PI * lum(constant);
            if (!fin(constant) || std::min({constant.x, constant.y, constant.z}) < 0)
                throw std::runtime_error("Invalid environment radiance");
            return;
        }
        if (d.width <= 0 || d.height <= 0) throw std::runtime_error("Invalid environment image dimensions");
        w = d.width;
        h = d.height;
        rot = std::fmod(d.rotation_degrees, 360.0f) * PI / 180;
        inten = d.intensity;
        if (!std::isfinite(d.rotation_degrees) || !std::isfinite(d.intensity) || d.intensity < 0)
            throw std::runtime_error("Invalid environment rotation/intensity");
        px.resize(size_t(w) * h);
        for (int y = 0; y < h; ++y) {
            const double om = omega(y);
            for (int x = 0; x < w; ++x) {
                const float* q = d.rgba + 4 * (size_t(y) * w + x);
                const V c{q[0], q[1], q[2]};
                if (!fin(c) || std::min({c.x, c.y, c.z}) < 0)
                    throw std::runtime_error("Non-finite or negative HDR environment pixel");
                px[size_t(y) * w + x] = c;
                const double l = lum(c) * inten;
                integral += l * om;
                wsum += (l + 1e-9) * om;
                cdf.push_back(wsum);
            }
        }
    }
    void uv(V d, float& u, float& v) const {
        u = (std::atan2(d.z, d.x) - rot) / (2 * PI) + 0.5f;
        u -= std::floor(u);
        v = std::acos(std::clamp(d.y, -1.0f, 1.0f)) / PI;
    }
    V radiance(V d) const {
        if (px.empty()) return constant;
        float u, v;
        uv(d, u, v);
        const float fx = u * float(w) - 0.5f, fy = v * float(h) - 0.5f;
        const int x = int(std::floor(fx)), y = int(std::floor(fy));
        const float tx = fx - float(x), ty = fy - float(y);
        auto at = [&](int ix, int iy) {
            ix = (ix % w + w) % w;
            iy = std::clamp(iy, 0, h - 1);
            return px[size_t(iy) * w + ix];
        };
        return lerp(lerp(at(x, y), at(x + 1, y), tx), lerp(at(x, y + 1), at(x + 1, y + 1), tx), ty) * inten;
    }
    V sample(Rng& g, float& pdf) const {
        if (px.empty()) {
            const float y = 1 - 2 * g.uni(), phi = 2 * PI * g.uni(), r = std::sqrt(std::max(0.0f, 1 - y * y));
            pdf = 1 / (4 * PI);
            return {r * std::cos(phi), y, r * std::sin(phi)};
        }
        const double pick = g.uni() * wsum;
        const size_t i = std::min(size_t(std::upper_bound(cdf.begin(), cdf.end(), pick) - cdf.begin()), cdf.size() - 1);
        const int x = int(i % size_t(w)), y = int(i / size_t(w));
        const float u = (float(x) + g.uni()) / float(w);
        const float c0 = std::cos(PI * float(y) / float(h)), c1 = std::cos(PI * float(y + 1) / float(h));
        const float ct = c0 + (c1 - c0) * g.uni(), st = std::sqrt(std::max(0.0f, 1 - ct * ct));
        const float phi = (u - 0.5f) * 2 * PI + rot;
        pdf = float((cdf[i] - (i ? cdf[i - 1] : 0)) / (wsum * omega(y)));
        return {st * std::cos(phi), ct, st * std::sin(phi)};
    }
    float pdf(V d) const {
        if (px.empty()) return 1 / (4 * PI);
        float u, v;
        uv(d, u, v);
        const int x = std::min(int(u * float(w)), w - 1), y = std::min(int(v * float(h)), h - 1);
        const size_t i = size_t(y) * w + x;
        return float((cdf[i] - (i ? cdf[i - 1] : 0)) / (wsum * omega(y)));
    }
};

// ---------------------------------------------------------------- BVH
struct Bvh {
    struct Node {
        V lo{INF, INF, INF}, hi{-INF, -INF, -INF};
        uint32_t first = 0, count = 0;
    };
    std::vector<Node> nodes;
    std::vector<uint32_t> idx;

    static V centroid(const Tri& t) {
        const D c = (dd(t.p[0]) + dd(t.p[1]) + dd(t.p[2])) * (1.0 / 3.0);
        return {float(c.x), float(c.y), float(c.z)};
    }
    static double area(V lo, V hi) {
        if (lo.x > hi.x || lo.y > hi.y || lo.z > hi.z) return 0;
        const D d = dd(hi) - dd(lo);
        return 2 * (d.x * d.y + d.y * d.z + d.z * d.x);
    }
    void build(const Scene& s) {
        nodes.clear();
        idx.resize(s.tris.size());
        std::iota(idx.begin(), idx.end(), 0u);
        if (s.tris.empty()) return;
        nodes.reserve(s.tris.size());
        nodes.emplace_back();
        split(s, 0, 0, uint32_t(s.tris.size()), 0);
    }
    void split(const Scene& s, uint32_t ni, uint32_t b, uint32_t e, int depth) {
        V lo{INF, INF, INF}, hi{-INF, -INF, -INF}, clo = lo, chi = hi;
        for (uint32_t i = b; i < e; ++i) {
            const Tri& t = s.tris[idx[i]];
            for (const V& p : t.p) {
                lo = vmin(lo, p);
                hi = vmax(hi, p);
            }
            const V c = centroid(t);
            clo = vmin(clo, c);
            chi = vmax(chi, c);
        }
        nodes[ni] = {lo, hi, b, e - b};
        if (e - b <= 4 || depth >= 48) return;
        const double pa = area(lo, hi);
        if (!(pa > 0)) return;
        struct Bin {
            V lo{INF, INF, INF}, hi{-INF, -INF, -INF};
            uint32_t n = 0;
        };
        std::array<std::array<Bin, 16>, 3> bins{};
        auto slot = [&](float c, int a) {
            const double ext = double(chi[a]) - clo[a];
            if (!(ext > 0)) return 0;
            return std::clamp(int((double(c) - clo[a]) / ext * 16), 0, 15);
        };
        for (uint32_t i = b; i < e; ++i) {
            const Tri& t = s.tris[idx[i]];
            V tl{INF, INF, INF}, th{-INF, -INF, -INF};
            for (const V& p : t.p) {
                tl = vmin(tl, p);
                th = vmax(th, p);
            }
            const V c = centroid(t);
            for (int a = 0; a < 3; ++a) {
                Bin& bn = bins[a][slot(c[a], a)];
                bn.lo = vmin(bn.lo, tl);
                bn.hi = vmax(bn.hi, th);
                ++bn.n;
            }
        }
        int ba = -1, bs = -1;
        double best = double(e - b) * pa;
        for (int a = 0; a < 3; ++a) {
            if (clo[a] == chi[a]) continue;
            std::array<V, 16> rlo, rhi;
            std::array<uint32_t, 16> rn{};
            V l{INF, INF, INF}, h{-INF, -INF, -INF};
            uint32_t n = 0;
            for (int k = 15; k >= 0; --k) {
                if (bins[a][k].n) {
                    l = vmin(l, bins[a][k].lo);
                    h = vmax(h, bins[a][k].hi);
                }
                n += bins[a][k].n;
                rlo[k] = l;
                rhi[k] = h;
                rn[k] = n;
            }
            l = {INF, INF, INF};
            h = {-INF, -INF, -INF};
            n = 0;
            for (int k = 0; k < 15; ++k) {
                if (bins[a][k].n) {
                    l = vmin(l, bins[a][k].lo);
                    h = vmax(h, bins[a][k].hi);
                }
                n += bins[a][k].n;
                if (!n || !rn[k + 1]) continue;
                const double cost = pa + n * area(l, h) + rn[k + 1] * area(rlo[k + 1], rhi[k + 1]);
                if (cost < best) {
                    best = cost;
                    ba = a;
                    bs = k;
                }
            }
        }
        if (ba < 0) return;
        const auto mid = std::partition(idx.begin() + b, idx.begin() + e,
                                        [&](uint32_t i) { return slot(centroid(s.tris[i])[ba], ba) <= bs; });
        const uint32_t m = uint32_t(mid - idx.begin());
        if (m == b || m == e) return;
        const uint32_t c0 = uint32_t(nodes.size());
        nodes.resize(nodes.size() + 2);
        nodes[ni].first = c0;
        nodes[ni].count = 0;
        split(s, c0, b, m, depth + 1);
        split(s, c0 + 1, m, e, depth + 1);
    }

    bool intersect(const Scene& s, const Ray& r, Hit& hit, bool any, uint32_t ignore = NONE) const {
        if (nodes.empty() || !validRay(r) || std::isnan(hit.t)) return false;
        Ray cl = r;
        double lim = std::min(double(r.tmax), hit.t);
        cl.tmax = std::min(cl.tmax, std::nextafter(float(lim), INF));
        const float qlim = cl.tmax;
        if (qlim < r.tmin) return false;
        double inv[3] = {0, 0, 0};
        const D o = dd(r.o) + dd(r.lo);
        for (int a = 0; a < 3; ++a)
            if (r.d[a] != 0) inv[a] = 1.0 / r.d[a];
        auto box = [&](const Node& n, float& nearOut) {
            double tn = cl.tmin, tf = cl.tmax;
            for (int a = 0; a < 3; ++a) {
                if (r.d[a] == 0) {
                    if (o[a] < n.lo[a] || o[a] > n.hi[a]) return false;
                    continue;
                }
                double t0 = (double(n.lo[a]) - o[a]) * inv[a], t1 = (double(n.hi[a]) - o[a]) * inv[a];
                if (inv[a] < 0) std::swap(t0, t1);
                const double eps = 4 * std::numeric_limits<double>::epsilon();
                t0 -= std::abs(t0) * eps;
                t1 += std::abs(t1) * eps;
                tn = std::max(tn, t0);
                tf = std::min(tf, t1);
                if (tn > tf) return false;
            }
            nearOut = std::max(cl.tmin, std::nextafter(float(tn), -INF));
            return true;
        };
        float rn = 0;
        if (!box(nodes[0], rn)) return false;
        struct E {
            uint32_t n;
            float near;
        };
        std::array<E, 50> st{};
        size_t sp = 0;
        st[sp++] = {0, rn};
        bool found = false;
        while (sp) {
            const E e = st[--sp];
            if (e.near > cl.tmax) continue;
            const Node& n = nodes[e.n];
            if (n.count) {
                for (uint32_t i = n.first; i < n.first + n.count; ++i) {
                    const uint32_t k = idx[i];
                    if (k == ignore) continue;
                    Hit c;
                    c.t = lim;
                    c.tri = k;
                    if (!hitTriangle(s.tris[k], cl, c) || !accepted(s, s.tris[k], cl, c)) continue;
                    if (any) {
                        hit = c;
                        return true;
                    }
                    if (!found || c.t < hit.t || (c.t == hit.t && k < hit.tri)) {
                        hit = c;
                        lim = c.t;
                        cl.tmax = std::min(qlim, std::nextafter(float(c.t), INF));
                        found = true;
                    }
                }
            } else {
                float ln = 0, rn2 = 0;
                const bool l = box(nodes[n.first], ln), rr = box(nodes[n.first + 1], rn2);
                if (sp + size_t(l) + size_t(rr) > st.size()) throw std::logic_error("BVH traversal exceeded its bounded stack");
                if (l && rr) {
                    if (ln <= rn2) {
                        st[sp++] = {n.first + 1, rn2};
                        st[sp++] = {n.first, ln};
                    } else {
                        st[sp++] = {n.first, ln};
                        st[sp++] = {n.first + 1, rn2};
                    }
                } else if (l) st[sp++] = {n.first, ln};
                else if (rr) st[sp++] = {n.first + 1, rn2};
            }
        }
        return found;
    }
};

// ---------------------------------------------------------------- lights
struct LightSmp {
    V dir, rad, target, tn;
    float dist = INF, pdf = 0;
    bool delta = false;
};
struct Lights {
    struct Entry {
        int type;
        uint32_t index;
        double cdf;
        float prob;
    };
    std::vector<Entry> es;
    std::vector<float> tprob;
    float envProb = 0;
    const Scene& s;
    const Env& env;
    Lights(const Scene& sc, const Env& e) : tprob(sc.tris.size()), s(sc), env(e) {
        V lo{INF, INF, INF}, hi{-INF, -INF, -INF};
        for (const Tri& t : sc.tris)
            for (const V& p : t.p) {
                lo = vmin(lo, p);
                hi = vmax(hi, p);
            }
        const double r2 = sc.tris.empty() ? 1 : std::max(1.0, double(len2(hi - lo)) * 0.25);
        double tot = 0;
        auto add = [&](int ty, uint32_t i, double w) {
            if (w > 0) {
                tot += w;
                es.push_back({ty, i, tot, 0});
            }
        };
        for (uint32_t i = 0; i < sc.lights.size(); ++i) {
            const Punctual& l = sc.lights[i];
            add(0, i, double(lum(l.color)) * l.intensity * (l.type == 1 ? PI * r2 : 4 * PI));
        }
        for (uint32_t i = 0; i < sc.tris.size(); ++i) {
            const Mat& m = sc.mats[sc.tris[i].mat];
            add(1, i, double(sc.tris[i].area) * PI * lum(m.emission) * (m.twoSided ? 2 : 1));
        }
        add(2, 0, env.integral * PI * r2);
        double last = 0, cum = 0;
        for (Entry& x : es) {
            const double p = 0.95 * (x.cdf - last) / tot + 0.05 / double(es.size());
            last = x.cdf;
            cum += p;
            x.prob = float(p);
            x.cdf = cum;
            if (x.type == 1) tprob[x.index] = x.prob;
            if (x.type == 2) envProb = x.prob;
        }
        if (!es.empty()) es.back().cdf = 1;
    }
    float envPdf(V d) const { return envProb * env.pdf(d); }
    float triPdf(uint32_t i, V from, V pt) const {
        const Tri& t = s.tris[i];
        const V d = pt - from;
        const float l2 = len2(d), c = std::abs(dot(t.ng, norm(d)));
        return c > 1e-8f ? tprob[i] * l2 / (t.area * c) : 0;
    }
    LightSmp sample(V from, Rng& g) const {
        LightSmp out;
        if (es.empty()) return out;
        const double u = g.uni();
        auto it = std::lower_bound(es.begin(), es.end(), u, [](const Entry& e, double v) { return e.cdf < v; });
        const Entry& e = it == es.end() ? es.back() : *it;
        if (e.type == 2) {
            out.dir = env.sample(g, out.pdf);
            out.pdf *= e.prob;
            out.rad = env.radiance(out.dir);
            return out;
        }
        if (e.type == 0) {
            const Punctual& l = s.lights[e.index];
            out.delta = true;
            out.pdf = e.prob;
            float att = 1;
            if (l.type == 1) out.dir = -l.dir;
            else {
                const V d = l.pos - from;
                const float r2 = len2(d);
                if (r2 < 1e-12f) return {};
                out.dist = std::sqrt(r2);
                out.dir = d / out.dist;
                out.target = l.pos;
                att = 1 / r2;
                if (std::isfinite(l.range)) {
                    const float q = out.dist / l.range;
                    att *= std::max(0.0f, 1 - q * q * q * q);
                }
                if (l.type == 2) {
                    const float ca = dot(l.dir, -out.dir), co = std::cos(l.outer), ci = std::cos(l.inner);
                    const float fo = std::clamp((ca - co) / std::max(1e-7f, ci - co), 0.0f, 1.0f);
                    att *= fo * fo;
                }
            }
            out.rad = l.color * (l.intensity * att);
            return out;
        }
        const Tri& t = s.tris[e.index];
        const float r = std::sqrt(g.uni()), u1 = 1 - r, u2 = r * g.uni(), u0 = 1 - u1 - u2;
        out.target = t.p[0] * u0 + t.p[1] * u1 + t.p[2] * u2;
        out.tn = t.ng;
        const V d = out.target - from;
        out.dist = len(d);
        if (out.dist < 1e-7f) return {};
        out.dir = d / out.dist;
        const Hit h{out.dist, u1, u2, e.index};
        const Ray ray{from, out.dir, 0, INF, {}};
        if (!accepted(s, t, ray, h)) return {};
        const Surf sf = surfaceAt(s, ray, h);
        if (!sf.front && !sf.twoSided) return {};
        out.rad = sf.emission;
        out.pdf = triPdf(e.index, from, out.target);
        return out;
    }
};

// ---------------------------------------------------------------- integrator
struct Boundary : std::runtime_error {
    using std::runtime_error::runtime_error;
};

struct Counters {
    uint64_t extend = 0, shadow = 0;
};

V trace(const Scene& sc, const Bvh& bvh, const Env& env, const Lights& lights, Ray ray, const crg_options& opt,
        Rng& g, bool precise, Counters& cnt) {
    V rad{}, thr{1, 1, 1}, lastPos{}, sigma{};
    float lastPdf = 0, etaScale = 1;
    bool lastDelta = true;
    int medium = -1;
    uint32_t prev = NONE;
    const bool nee = opt.next_event_estimation != 0;
    for (int depth = 0; depth <= opt.bounces; ++depth) {
        Hit hit;
        ++cnt.extend;
        if (!bvh.intersect(sc, ray, hit, false, prev)) {
            if (medium >= 0) thr = thr * beer(sigma, INF);
            const float w = lastDelta || !nee ? 1 : mis(lastPdf, lights.envPdf(ray.d));
            rad = rad + thr * env.radiance(ray.d) * w;
            break;
        }
        if (medium >= 0) thr = thr * beer(sigma, float(hit.t));
        const Surf s = surfaceAt(sc, ray, hit);
        if (s.front || s.twoSided) {
            const float w = lastDelta || !nee ? 1 : mis(lastPdf, lights.triPdf(hit.tri, lastPos, s.p));
            rad = rad + thr * s.emission * w;
        }
        if (depth == opt.bounces) break;
        const V wo = -ray.d;
        if (nee) {
            const LightSmp l = lights.sample(s.p, g);
            const bool enclosedInf = medium >= 0 && s.thickness <= 0 && !std::isfinite(l.dist);
            if (!enclosedInf && l.pdf > 0 && maxc(l.rad) > 0) {
                const Eval b = evalBsdf(s, wo, l.dir);
                if (maxc(b.value) > 0) {
                    Ray sh = precise ? Ray{s.p, l.dir, 0, INF, s.plo} : offset(s.p, s.ng, l.dir, INF);
                    if (std::isfinite(l.dist)) {
                        V end = l.target;
                        if (len2(l.tn) > 0) end = offset(end, l.tn, -l.dir, INF).o;
                        const V d = end - sh.o;
                        sh.tmax = len(d) * (1 - 2e-6f);
                        sh.d = norm(d);
                    }
                    Hit blocker;
                    ++cnt.shadow;
                    if (!bvh.intersect(sc, sh, blocker, true, precise ? hit.tri : NONE)) {
                        const float w = l.delta ? 1 : mis(l.pdf, b.pdf);
                        V shSigma = sigma;
                        bool shIn = medium >= 0;
                        if (s.thickness > 0 && dot(wo, s.ng) * dot(l.dir, s.ng) < 0) {
                            shIn = s.front;
                            shSigma = s.front ? absorb(s.attColor, s.attDist) : V{};
                        }
                        const V tr = shIn ? beer(shSigma, l.dist) : V{1, 1, 1};
                        rad = rad + thr * b.value * l.rad * tr * (std::abs(dot(s.n, l.dir)) * w / l.pdf);
                    }
                }
            }
        }
        const Smp smp = sampleBsdf(s, wo, g);
        if (!smp.ok()) break;
        thr = thr * smp.w;
        if (!fin(thr)) throw std::runtime_error("Non-finite path throughput");
        if (smp.tr && s.thickness > 0) {
            if (s.front) {
                if (medium >= 0)
                    throw Boundary("Nested/intersecting refractive volumes are unsupported; separate the closed meshes");
                medium = int(s.mat);
                sigma = absorb(s.attColor, s.attDist);
            } else {
                if (medium < 0) {
                    std::ostringstream m;
                    m.precision(9);
                    m << "Refractive exit without entry; check outward normals and keep the camera outside closed "
                         "volumes; bounce "
                      << depth << ", triangle " << hit.tri << ", material " << s.mat << ", t " << hit.t << ", point ("
                      << s.p.x << ',' << s.p.y << ',' << s.p.z << "), origin (" << ray.o.x << ',' << ray.o.y << ','
                      << ray.o.z << "), direction (" << ray.d.x << ',' << ray.d.y << ',' << ray.d.z << ')';
                    throw Boundary(m.str());
                }
                medium = -1;
                sigma = {};
            }
            etaScale *= smp.eta * smp.eta;
        }
        lastPdf = smp.pdf;
        lastDelta = smp.delta;
        lastPos = s.p;
        ray = precise ? Ray{s.p, smp.dir, 0, INF, s.plo} : offset(s.p, s.ng, smp.dir, INF);
        prev = precise ? hit.tri : NONE;
        if (opt.russian_roulette && depth >= 3) {
            const float surv = std::clamp(maxc(thr) * etaScale, 0.05f, 0.95f);
            if (g.uni() >= surv) break;
            thr = thr * (1 / surv);
        }
    }
    if (!fin(rad) || std::min({rad.x, rad.y, rad.z}) < 0) throw std::runtime_error("Non-finite or negative path radiance");
    return rad;
}

struct RenderOut {
    uint64_t retries = 0, extend = 0, shadow = 0;
    int workers = 0;
    double seconds = 0;
};

void validate(const crg_options& o) {
    if (o.width < 1 || o.height < 1 || o.width > 16384 || o.height > 16384)
        throw std::runtime_error("Width and height must be in [1, 16384]");
    if (uint64_t(o.width) * uint64_t(o.height) > 67108864)
        throw std::runtime_error("Image exceeds the 64-megapixel allocation limit");
    if (o.samples < 1 || o.samples > 1048576) throw std::runtime_error("Samples must be in [1, 1048576]");
    if (o.bounces < 1 || o.bounces > 64) throw std::runtime_error("Bounces must be in [1, 64]");
    if (o.threads < 0 || o.threads > 256) throw std::runtime_error("Threads must be 0 (automatic) or in [1, 256]");
}

// A window (wx, wy, ww, wh) of the frame renders only those pixels, each with
// its global index and hence its full-frame sample stream; film (and, when
// given, samples: [sample][window pixel][rgb], the radiance of every sample in
// order) is indexed by window pixel.  The full frame is the window (0, 0, W, H).
RenderOut render(const Scene& sc, const Bvh& bvh, const Cam& cam, const Env& env, const crg_options& o, float* film,
                 int wx = 0, int wy = 0, int ww = -1, int wh = -1, float* samples = nullptr) {
    validate(o);
    if (ww < 0) ww = o.width;
    if (wh < 0) wh = o.height;
    if (wx < 0 || wy < 0 || ww < 1 || wh < 1 || wx + ww > o.width || wy + wh > o.height)
        throw std::runtime_error("Render window must lie inside the image");
    RenderOut out;
    const int tile = 16, tx = (ww + tile - 1) / tile, ty = (wh + tile - 1) / tile, tiles = tx * ty;
    out.workers = std::min(tiles, o.threads ? o.threads : int(std::max(1u, std::thread::hardware_concurrency() / 2)));
    const Lights lights(sc, env);
    std::atomic<int> next{0};
    std::atomic<uint64_t> retries{0}, ext{0}, shd{0};
    std::atomic<bool> stop{false};
    std::mutex mu;
    std::string err;
    const auto t0 = std::chrono::steady_clock::now();
    auto work = [&] {
        int cx = -1, cy = -1, cs = -1;
        Counters cnt;
        try {
            for (;;) {
                const int t = next.fetch_add(1);
                if (t >= tiles || stop.load()) break;
                const int x0 = wx + (t % tx) * tile, y0 = wy + (t / tx) * tile;
                const int x1 = std::min(x0 + tile, wx + ww), y1 = std::min(y0 + tile, wy + wh);
                for (int y = y0; y < y1; ++y)
                    for (int x = x0; x < x1; ++x) {
                        cx = x;
                        cy = y;
                        const uint64_t pixel = uint64_t(y) * o.width + x;
                        const uint64_t local = uint64_t(y - wy) * ww + uint64_t(x - wx);
                        double r = 0, gsum = 0, b = 0;
                        for (int smp = 0; smp < o.samples; ++smp) {
                            cs = smp;
                            Rng g{seedOf(o.seed, pixel, uint64_t(smp))};
                            const float dx = g.uni(), dy = g.uni();
                            const Ray ray = camRay(cam, float(x) + dx, float(y) + dy, o.width, o.height);
                            const Rng first = g;
                            V l;
                            try {
                                l = trace(sc, bvh, env, lights, ray, o, g, false, cnt);
                            } catch (const Boundary&) {
                                retries.fetch_add(1);
                                g = first;
                                l = trace(sc, bvh, env, lights, ray, o, g, true, cnt);
                            }
                            r += l.x;
                            gsum += l.y;
                            b += l.z;
                            if (samples) {
                                float* q = samples + 3 * (uint64_t(smp) * uint64_t(ww) * uint64_t(wh) + local);
                                q[0] = l.x;
                                q[1] = l.y;
                                q[2] = l.z;
                            }
                        }
                        film[3 * local] = float(r / o.samples);
                        film[3 * local + 1] = float(gsum / o.samples);
                        film[3 * local + 2] = float(b / o.samples);
                    }
            }
        } catch (const std::exception& e) {
            stop.store(true);
            std::lock_guard<std::mutex> lk(mu);
            if (err.empty())
                err = "Pixel (" + std::to_string(cx) + "," + std::to_string(cy) + "), sample " + std::to_string(cs) +
                      ": " + e.what();
        }
        ext.fetch_add(cnt.extend);
        shd.fetch_add(cnt.shadow);
    };
    std::vector<std::thread> pool;
    for (int i = 0; i < out.workers; ++i) pool.emplace_back(work);
    for (auto& t : pool) t.join();
    if (!err.empty()) throw std::runtime_error(err);
    out.retries = retries.load();
    out.extend = ext.load();
    out.shadow = shd.load();
    out.seconds = std::chrono::duration<double>(std::chrono::steady_clock::now() - t0).count();
    return out;
}

// ---------------------------------------------------------------- probe glue
thread_local std::string lastError;

struct SceneBox {
    Scene s;
    Bvh bvh;
};

Cam camOf(const crg_camera& c) {
    Cam k;
    k.pos = v3(c.position);
    k.fwd = v3(c.forward);
    k.right = v3(c.right);
    k.up = v3(c.up);
    k.yfov = c.yfov;
    k.aspect = c.aspect;
    k.nearc = c.near_clip;
    k.farc = c.far_clip;
    return k;
}
Ray rayOf(const crg_ray& r) { return {v3(r.origin), v3(r.direction), r.tmin, r.tmax, v3(r.origin_low)}; }
void putRay(crg_ray* o, const Ray& r) {
    put(o->origin, r.o);
    put(o->direction, r.d);
    o->tmin = r.tmin;
    o->tmax = r.tmax;
    put(o->origin_low, r.lo);
    o->ignore = NONE;
}
Surf surfOf(const probe_surface& p) {
    Surf s;
    s.p = v3(p.position);
    s.plo = v3(p.position_low);
    s.ng = v3(p.geometric_normal);
    s.n = v3(p.normal);
    s.tan = v3(p.tangent);
    s.bit = v3(p.bitangent);
    s.base = v3(p.base_color);
    s.emission = v3(p.emission);
    s.attColor = v3(p.attenuation_color);
    s.metallic = p.metallic;
    s.roughness = p.roughness;
    s.transmission = p.transmission;
    s.ior = p.ior;
    s.thickness = p.thickness;
    s.attDist = p.attenuation_distance;
    s.mat = p.material;
    s.front = p.front_face != 0;
    s.twoSided = p.double_sided != 0;
    return s;
}
void putSurf(probe_surface* o, const Surf& s) {
    put(o->position, s.p);
    put(o->position_low, s.plo);
    put(o->geometric_normal, s.ng);
    put(o->normal, s.n);
    put(o->tangent, s.tan);
    put(o->bitangent, s.bit);
    put(o->base_color, s.base);
    put(o->emission, s.emission);
    put(o->attenuation_color, s.attColor);
    o->metallic = s.metallic;
    o->roughness = s.roughness;
    o->transmission = s.transmission;
    o->ior = s.ior;
    o->thickness = s.thickness;
    o->attenuation_distance = s.attDist;
    o->material = s.mat;
    o->front_face = s.front;
    o->double_sided = s.twoSided;
}

template <class F>
int guard(F&& f) {
    try {
        f();
        return 0;
    } catch (const std::exception& e) {
        lastError = e.what();
        return 1;
    }
}

}  // namespace orc

using namespace orc;

namespace orc {
// ---------------------------------------------------------------- film encode
void checkOutput(int w, int h, const float* film, const std::string& label, bool hdr) {
    if (w <= 0 || h <= 0 || w > 16384 || h > 16384 || size_t(w) > (16ull * 1024 * 1024) / size_t(h))
        throw std::runtime_error(label + ": image dimensions must be positive, at most 16384 per axis and 16777216 pixels");
    const size_t n = size_t(w) * size_t(h);
    for (size_t i = 0; i < n; ++i) {
        const float* p = film + 3 * i;
        if (!std::isfinite(p[0]) || !std::isfinite(p[1]) || !std::isfinite(p[2]) || p[0] < 0 || p[1] < 0 || p[2] < 0)
            throw std::runtime_error(label + ": pixels must contain finite, nonnegative linear radiance");
        if (hdr && std::max({p[0], p[1], p[2]}) >= std::ldexp(1.0f, 127))
            throw std::runtime_error(label + ": radiance exceeds the Radiance RGBE exponent range");
    }
}
// Filmic (ACES-like rational), sRGB, 8-bit (reference src/io/image.cpp:117-124, src/math.h:62).
float toSrgb(float v) { return v <= 0.0031308f ? 12.92f * v : 1.055f * std::pow(v, 1 / 2.4f) - 0.055f; }
uint8_t displayByte(float value, double scale) {
    const double x = double(value) * scale;
    const double f = std::clamp((x * (2.51 * x + 0.03)) / (x * (2.43 * x + 0.59) + 0.14), 0.0, 1.0);
    const float s = toSrgb(float(f));
    return uint8_t(std::lround(std::clamp(s, 0.0f, 1.0f) * 255.0f));
}
// Shared-exponent RGBE of one pixel.
void toRgbe(const float* c, uint8_t out[4]) {
    const float m = c[0] > (c[1] > c[2] ? c[1] : c[2]) ? c[0] : (c[1] > c[2] ? c[1] : c[2]);
    if (m < 1e-32f) {
        out[0] = out[1] = out[2] = out[3] = 0;
        return;
    }
    int e;
    const float norm = float(std::frexp(m, &e)) * 256.0f / m;
    for (int k = 0; k < 3; ++k) out[k] = uint8_t(c[k] * norm);
    out[3] = uint8_t(e + 128);
}
// New-style run-length encoding of one component of a scanline: literal
// blocks of at most 128 bytes up to the next run of three equal bytes, runs in
// blocks of at most 127.
void rleComponent(const uint8_t* c, int w, std::vector<uint8_t>& out) {
    int x = 0;
    while (x < w) {
        int r = x;
        while (r + 2 < w && !(c[r] == c[r + 1] && c[r] == c[r + 2])) ++r;
        const bool run = r + 2 < w;
        if (!run) r = w;
        while (x < r) {
            const int len = std::min(r - x, 128);
            out.push_back(uint8_t(len));
            out.insert(out.end(), c + x, c + x + len);
            x += len;
        }
        if (run) {
            while (r < w && c[r] == c[x]) ++r;
            while (x < r) {
                const int len = std::min(r - x, 127);
                out.push_back(uint8_t(128 + len));
                out.push_back(c[x]);
                x += len;
            }
        }
    }
}
std::vector<uint8_t> hdrFile(int w, int h, const float* film) {
    std::vector<uint8_t> out;
    const std::string head = "#?RADIANCE\n# Written by stb_image_write.h\nFORMAT=32-bit_rle_rgbe\n"
                             "EXPOSURE=          1.0000000000000\n\n-Y " + std::to_string(h) + " +X " + std::to_string(w) + "\n";
    out.insert(out.end(), head.begin(), head.end());
    std::vector<uint8_t> comp(size_t(w) * 4);
    for (int y = 0; y < h; ++y) {
        const float* row = film + size_t(y) * size_t(w) * 3;
        if (w < 8 || w >= 32768) {
            for (int x = 0; x < w; ++x) {
                uint8_t q[4];
                toRgbe(row + 3 * x, q);
                out.insert(out.end(), q, q + 4);
            }
            continue;
        }
        for (int x = 0; x < w; ++x) {
            uint8_t q[4];
            toRgbe(row + 3 * x, q);
            for (int k = 0; k < 4; ++k) comp[size_t(k) * w + x] = q[k];
        }
        const uint8_t sh[4] = {2, 2, uint8_t((w >> 8) & 0xff), uint8_t(w & 0xff)};
        out.insert(out.end(), sh, sh + 4);
        for (int k = 0; k < 4; ++k) rleComponent(comp.data() + size_t(k) * w, w, out);
    }
    return out;
}
}  // namespace orc

extern "C" {

const char* orc_last_error(void) { return lastError.c_str(); }
uint64_t orc_sample_seed(uint64_t seed, uint64_t pixel, uint64_t sample) { return seedOf(seed, pixel, sample); }
void orc_rng_draw(uint64_t* state, int count, float* u) {
    Rng g{*state};
    for (int i = 0; i < count; ++i) u[i] = g.uni();
    *state = g.s;
}
int orc_camera_generate(const crg_camera* c, float px, float py, int w, int h, crg_ray* out) {
    return guard([&] { putRay(out, camRay(camOf(*c), px, py, w, h)); });
}

void* orc_scene_create(const crg_scene_desc* d) {
    auto box = std::make_unique<SceneBox>();
    const int rc = guard([&] {
        Scene& s = box->s;
        for (uint32_t i = 0; i < d->material_count; ++i) {
            const crg_material& m = d->materials[i];
            Mat x;
            auto slotOf = [](const crg_texture_slot& t) {
                Slot o;
                o.texture = t.texture;
                o.texcoord = t.texcoord;
                o.offset[0] = t.offset[0];
                o.offset[1] = t.offset[1];
                o.scale[0] = t.scale[0];
                o.scale[1] = t.scale[1];
                o.rotation = t.rotation;
                return o;
            };
            x.baseTex = slotOf(m.base_color_tex);
            x.mrTex = slotOf(m.metallic_roughness_tex);
            x.normalTex = slotOf(m.normal_tex);
            x.emissiveTex = slotOf(m.emissive_tex);
            x.transmissionTex = slotOf(m.transmission_tex);
            x.normalScale = m.normal_scale;
            x.base = v3(m.base_color);
            x.alpha = m.base_color[3];
            x.metallic = m.metallic;
            x.roughness = m.roughness;
            x.emission = v3(m.emission);
            x.transmission = m.transmission;
            x.ior = m.ior;
            x.thickness = m.thickness;
            x.attDist = m.attenuation_distance;
            x.attColor = v3(m.attenuation_color);
            x.twoSided = m.double_sided != 0;
            x.mask = m.alpha_mode == 1;
            x.cutoff = m.alpha_cutoff;
            s.mats.push_back(x);
        }
        for (uint32_t i = 0; i < d->light_count; ++i) {
            const crg_light& l = d->lights[i];
            Punctual p;
            p.type = l.type;
            p.pos = v3(l.position);
            p.dir = v3(l.direction);
            p.color = v3(l.color);
            p.intensity = l.intensity;
            p.range = l.range;
            p.inner = l.inner_cone;
            p.outer = l.outer_cone;
            s.lights.push_back(p);
        }
        for (uint32_t i = 0; i < d->image_count; ++i) {
            Image im;
            im.w = d->images[i].width;
            im.h = d->images[i].height;
            if (im.w <= 0 || im.h <= 0 || !d->images[i].rgba)
                throw std::runtime_error("Image " + std::to_string(i) + " has invalid dimensions or pixel count");
            im.px.resize(size_t(im.w) * size_t(im.h));
            for (size_t q = 0; q < im.px.size(); ++q)
                for (int k = 0; k < 4; ++k) im.px[q][k] = d->images[i].rgba[4 * q + k];
            s.images.push_back(std::move(im));
        }
        for (uint32_t i = 0; i < d->texture_count; ++i) {
            Texture t;
            t.image = d->textures[i].image;
            t.wrapS = d->textures[i].wrap_s;
            t.wrapT = d->textures[i].wrap_t;
            t.linear = d->textures[i].linear != 0;
            if (t.image < 0 || size_t(t.image) >= s.images.size())
                throw std::runtime_error("Texture " + std::to_string(i) + " has an invalid image index");
            s.textures.push_back(t);
        }
        for (size_t i = 0; i < s.mats.size(); ++i)
            for (const Slot* sl : {&s.mats[i].baseTex, &s.mats[i].mrTex, &s.mats[i].normalTex, &s.mats[i].emissiveTex,
                                   &s.mats[i].transmissionTex})
                if (sl->texture < -1 || (sl->texture >= 0 && size_t(sl->texture) >= s.textures.size()))
                    throw std::runtime_error("Material " + std::to_string(i) + " has an invalid texture index");
        s.tris.resize(d->triangle_count);
        for (uint64_t t = 0; t < d->triangle_count; ++t) {
            Tri& tr = s.tris[t];
            tr.mat = d->material_ids[t];
            for (int k = 0; k < 3; ++k) {
                tr.p[k] = v3(d->positions + 9 * t + 3 * k);
                tr.n[k] = d->normals ? v3(d->normals + 9 * t + 3 * k) : V{};
                for (int q = 0; q < 4; ++q) tr.tan[k][q] = d->tangents ? d->tangents[12 * t + 4 * k + q] : (q == 3 ? 1.0f : 0.0f);
                for (int q = 0; q < 4; ++q) tr.uv[k][q / 2][q % 2] = d->uvs ? d->uvs[12 * t + 4 * k + q] : 0.0f;
                if (!fin(tr.p[k]) || !fin(tr.n[k]))
                    throw std::runtime_error("Triangle " + std::to_string(t) + " contains non-finite vertex data");
            }
        }
        finalize(s);
        box->bvh.build(s);
    });
    return rc ? nullptr : box.release();
}
void orc_scene_destroy(void* s) { delete static_cast<SceneBox*>(s); }
uint64_t orc_scene_triangles(void* s) { return static_cast<SceneBox*>(s)->s.tris.size(); }
uint64_t orc_scene_nodes(void* s) { return static_cast<SceneBox*>(s)->bvh.nodes.size(); }

int orc_trace(void* scene, const crg_ray* rays, uint64_t n, crg_hit* hits, int any) {
    const SceneBox& b = *static_cast<SceneBox*>(scene);
    return guard([&] {
        for (uint64_t i = 0; i < n; ++i) {
            Hit h;
            const bool f = b.bvh.intersect(b.s, rayOf(rays[i]), h, any != 0, rays[i].ignore);
            hits[i].hit = f;
            hits[i].triangle = f ? h.tri : NONE;
            hits[i].t = f ? h.t : 0;
            hits[i].u = f ? h.u : 0;
            hits[i].v = f ? h.v : 0;
        }
    });
}
int orc_offset_ray(const float* p, const float* n, const float* d, float maxd, crg_ray* out) {
    return guard([&] { putRay(out, offset(v3(p), v3(n), v3(d), maxd)); });
}
int orc_surface_at(void* scene, const crg_ray* ray, const crg_hit* hit, probe_surface* out) {
    const SceneBox& b = *static_cast<SceneBox*>(scene);
    return guard([&] {
        Hit h{hit->t, hit->u, hit->v, hit->triangle};
        putSurf(out, surfaceAt(b.s, rayOf(*ray), h));
    });
}
void orc_bsdf_sample(const probe_surface* s, const float* wo, uint64_t* rng, probe_bsdf_sample* out) {
    Rng g{*rng};
    const Smp b = sampleBsdf(surfOf(*s), v3(wo), g);
    *rng = g.s;
    put(out->direction, b.dir);
    put(out->weight, b.w);
    out->pdf = b.pdf;
    out->eta = b.eta;
    out->delta = b.delta;
    out->transmitted = b.tr;
    out->valid = b.ok();
}
void orc_bsdf_eval(const probe_surface* s, const float* wo, const float* wi, float* vp) {
    const Eval e = evalBsdf(surfOf(*s), v3(wo), v3(wi));
    put(vp, e.value);
    vp[3] = e.pdf;
}
void* orc_env_create(const crg_env_desc* d) {
    auto e = std::make_unique<Env>();
    if (guard([&] { e->init(*d); })) return nullptr;
    return e.release();
}
void orc_env_destroy(void* e) { delete static_cast<Env*>(e); }
void orc_env_sample(void* e, uint64_t* rng, float* dir, float* pdf) {
    Rng g{*rng};
    put(dir, static_cast<Env*>(e)->sample(g, *pdf));
    *rng = g.s;
}
float orc_env_pdf(void* e, const float* d) { return static_cast<Env*>(e)->pdf(v3(d)); }
void orc_env_radiance(void* e, const float* d, float* rgb) { put(rgb, static_cast<Env*>(e)->radiance(v3(d))); }
int orc_render(void* scene, void* env, const crg_camera* c, const crg_options* o, float* film, crg_stats* st) {
    const SceneBox& b = *static_cast<SceneBox*>(scene);
    return guard([&] {
        const RenderOut r = render(b.s, b.bvh, camOf(*c), *static_cast<Env*>(env), *o, film);
        if (st) {
            std::memset(st, 0, sizeof(*st));
            st->seconds = r.seconds;
            st->camera_samples = uint64_t(o->width) * o->height * o->samples;
            st->boundary_retries = r.retries;
            st->extend_rays = r.extend;
            st->shadow_rays = r.shadow;
            st->workers = r.workers;
        }
    });
}

int orc_render_window(void* scene, void* env, const crg_camera* c, const crg_options* o, int x0, int y0, int w, int h,
                      float* film, float* samples, crg_stats* st) {
    const SceneBox& b = *static_cast<SceneBox*>(scene);
    return guard([&] {
        const RenderOut r = render(b.s, b.bvh, camOf(*c), *static_cast<Env*>(env), *o, film, x0, y0, w, h, samples);
        if (st) {
            std::memset(st, 0, sizeof(*st));
            st->seconds = r.seconds;
            st->camera_samples = uint64_t(w) * uint64_t(h) * uint64_t(o->samples);
            st->boundary_retries = r.retries;
            st->extend_rays = r.extend;
            st->shadow_rays = r.shadow;
            st->workers = r.workers;
        }
    });
}

int orc_film_png8(const float* film, int width, int height, float exposure, uint8_t* rgb8) {
    return guard([&] {
        orc::checkOutput(width, height, film, "PNG output", false);
        if (!std::isfinite(exposure) || exposure < -100 || exposure > 100)
            throw std::runtime_error("PNG exposure must be finite and between -100 and 100 stops");
        const double scale = std::exp2(double(exposure));
        const size_t n = size_t(width) * size_t(height) * 3;
        for (size_t i = 0; i < n; ++i) rgb8[i] = orc::displayByte(film[i], scale);
    });
}
int orc_film_hdr(const float* film, int width, int height, uint8_t* out, uint64_t capacity, uint64_t* size) {
    return guard([&] {
        orc::checkOutput(width, height, film, "HDR output", true);
        const std::vector<uint8_t> b = orc::hdrFile(width, height, film);
        *size = b.size();
        if (b.size() > capacity) throw std::runtime_error("HDR output buffer too small");
        std::memcpy(out, b.data(), b.size());
    });
}
}
