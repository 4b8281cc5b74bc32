// TEST INFRASTRUCTURE: CPU emulation of the gfx950 path.
//
// Compiles the device-side sources (path.h, geom.h, shade.h, sf64.h) and the
// host scene preparation (host_scene.cpp) with a plain C++ compiler and runs
// the exact per-path functions the kernels run: begin_sample -> [traverse
// closest -> bounce -> traverse any for the shadow ray] ... -> accumulate in
// emulated binary64.  Used by the CPU test-suite to check the device logic
// against the oracle without a GPU, and to reproduce device behaviour on the
// host when a GPU result needs explaining.
#include <array>
#include <map>
#include <algorithm>
#include <cmath>
#include <random>
#include <atomic>
#include <cstring>
#include <memory>
#include <mutex>
#include <string>
#include <thread>
#include <vector>

#include "../../cpu-renderer_amd/csrc/host_scene.h"
#include "../../cpu-renderer_amd/csrc/film.h"
#include "../../cpu-renderer_amd/csrc/film_host.h"
#include "../../cpu-renderer_amd/csrc/path.h"

using namespace cr;

namespace {

thread_local std::string lastError;

// The sampling-angle table, filled by the host code the library uses.
const bool kTurnTableReady = [] {
    static std::vector<TurnCS> table(kTurnEntries);
    fill_turn_table(table.data());
    g_turn_table = table.data();
    return true;
}();

struct EmuScene {
    HostScene hs;
    SceneArrays arr;
    std::vector<uint32_t> ranks;
    SceneLights lights;
    DevScene dev{};
};

struct EmuEnv {
    HostEnv he;
    std::vector<float4> texels;
    DevEnv dev{};
};

}  // namespace

// FP32 pass of the postponed-leaf traversal (inner()/leaves()) with the leaf
// rounds placed by a pseudo-random schedule: any interleaving must give the
// same result as step().  Odd schedules use the plain leaf test, even ones the
// preloading one.
template <bool ANY, bool PRELOAD, bool MASK = true, int W = 4>
ClosestResult phased(const DevScene& s, RayQuery& q, TraceCount& tc, uint64_t& rng, Pending* cands) {
    StackEntry overflow[kStackSize];
    uint32_t ring_ref[LocalRing::kSize];
    float ring_near[LocalRing::kSize];
    Traversal<ANY, false, LocalRing, PRELOAD, MASK, W> tr;
    tr.stack = overflow;
    tr.ring = LocalRing{ring_ref, ring_near};
    tr.cands = cands;
    if (tr.begin(s, q)) return tr.best;
    for (;;) {
        if (tr.inner(s, q, tc)) break;
        rng = rng * 6364136223846793005ull + 1442695040888963407ull;
        const bool go = !tr.can_descend() || ((rng >> 61) & 1);
        if (go && tr.has_leaf() && tr.leaves(s, q, tc)) break;
    }
    return tr.best;
}

// 1: a leaf's float box leaves the decoded boxes on its path; 2: the leaf set
// differs from the float tree's.
template <int W>
int check_quantized(const std::vector<QNode<W>>& q, const std::map<uint32_t, std::array<float, 6>>& leaf_box) {
    if (q.empty()) return leaf_box.empty() ? 0 : 2;
    struct Item {
        uint32_t ref;
        float lo[3], hi[3];
    };
    std::vector<Item> stack{{0, {-INFINITY, -INFINITY, -INFINITY}, {INFINITY, INFINITY, INFINITY}}};
    size_t leaves = 0;
    while (!stack.empty()) {
        const Item it = stack.back();
        stack.pop_back();
        if (ref_is_leaf(it.ref)) {
            auto f = leaf_box.find(it.ref);
            if (f == leaf_box.end()) return 2;
            ++leaves;
            for (int a = 0; a < 3; ++a)
                if (f->second[a] < it.lo[a] || f->second[3 + a] > it.hi[a]) return 1;
            continue;
        }
        if (it.ref >= q.size()) return 2;
        const QNode<W>& nd = q[it.ref];
        const float p[3] = {nd.px, nd.py, nd.pz};
        for (int c = 0; c < int(nd.exps >> 24); ++c) {
            Item ch{nd.child[c], {}, {}};
            for (int a = 0; a < 3; ++a) {
                const float sc = qscale((nd.exps >> (8 * a)) & 255u);
                const uint32_t ql = (nd.qlo[a][c / 4] >> (8 * (c % 4))) & 255u;
                const uint32_t qh = (nd.qhi[a][c / 4] >> (8 * (c % 4))) & 255u;
                ch.lo[a] = std::max(it.lo[a], qdecode(p[a], ql, sc));
                ch.hi[a] = std::min(it.hi[a], qdecode(p[a], qh, sc));
            }
            stack.push_back(ch);
        }
    }
    return leaves == leaf_box.size() ? 0 : 2;
}

extern "C" {

const char* emu_last_error(void) { return lastError.c_str(); }

void* emu_scene_create(const crg_scene_desc* d) {
    try {
        auto s = std::make_unique<EmuScene>();
        s->hs = finalize_scene(*d);
        s->arr = scene_arrays(s->hs);
        const size_t n = s->hs.tris.size();
        s->lights = scene_lights(s->hs);
        s->ranks = light_ranks(s->lights, s->arr.bvh.order, n);
        SceneArrays& a = s->arr;
        s->dev = DevScene{a.bvh.q4.data(), a.bvh.q8.data(), a.prims.data(), a.prim_ng.data(), a.aux.data(), a.materials.data(),
                          s->ranks.data(), s->lights.punctual.data(), a.uvs.empty() ? nullptr : a.uvs.data(),
                          a.tangents.empty() ? nullptr : a.tangents.data(), a.texels.data(), a.images.data(),
                          a.textures.data(), uint32_t(a.bvh.q4.size()), uint32_t(a.prims.size()),
                          uint32_t(s->lights.punctual.size()), uint32_t(a.bvh.q8.size())};
        return s.release();
    } catch (const std::exception& e) {
        lastError = e.what();
        return nullptr;
    }
}

void emu_scene_destroy(void* s) { delete static_cast<EmuScene*>(s); }
uint64_t emu_scene_triangles(void* s) { return static_cast<EmuScene*>(s)->hs.tris.size(); }
uint64_t emu_scene_wide_nodes(void* s) { return static_cast<EmuScene*>(s)->arr.bvh.nodes.size(); }

// Structural check of the 4-wide tree: every primitive in exactly one leaf,
// every child box encloses its subtree, and leaves respect the encoding.
int emu_scene_check(void* sp) {
    const EmuScene& s = *static_cast<EmuScene*>(sp);
    const auto& nodes = s.arr.bvh.nodes;
    const size_t n = s.arr.prims.size();
    if (n == 0) return nodes.empty() ? 0 : 1;
    std::vector<int> seen(n, 0);
    std::vector<uint32_t> refs(s.hs.tris.size(), 0);  // leaf-order primitives per triangle
    for (uint32_t t : s.arr.bvh.order) ++refs[t];
    for (uint32_t c : refs)
        if (c == 0) return 7;  // every triangle is referenced
    struct Item {
        uint32_t ref;
        float lo[3], hi[3];
    };
    std::vector<Item> stack{{0, {-INFINITY, -INFINITY, -INFINITY}, {INFINITY, INFINITY, INFINITY}}};
    while (!stack.empty()) {
        const Item it = stack.back();
        stack.pop_back();
        if (ref_is_leaf(it.ref)) {
            for (uint32_t i = leaf_first(it.ref); i < leaf_first(it.ref) + leaf_count(it.ref); ++i) {
                if (i >= n) return 2;
                ++seen[i];
                const PrimRec& p = s.arr.prims[i];
                const float xs[3] = {p.ax, p.bx, p.cx}, ys[3] = {p.ay, p.by, p.cy}, zs[3] = {p.az, p.bz, p.cz};
                if (refs[p.id] == 1) {  // an unsplit triangle lies inside its leaf box
                    for (int k = 0; k < 3; ++k)
                        if (xs[k] < it.lo[0] || xs[k] > it.hi[0] || ys[k] < it.lo[1] || ys[k] > it.hi[1] ||
                            zs[k] < it.lo[2] || zs[k] > it.hi[2])
                            return 3;
                } else {  // a split reference's box meets the triangle's box
                    const float lo[3] = {std::min({xs[0], xs[1], xs[2]}), std::min({ys[0], ys[1], ys[2]}),
                                         std::min({zs[0], zs[1], zs[2]})};
                    const float hi[3] = {std::max({xs[0], xs[1], xs[2]}), std::max({ys[0], ys[1], ys[2]}),
                                         std::max({zs[0], zs[1], zs[2]})};
                    for (int a = 0; a < 3; ++a)
                        if (hi[a] < it.lo[a] || lo[a] > it.hi[a]) return 3;
                }
            }
            continue;
        }
        if (it.ref >= nodes.size()) return 4;
        const WNode& w = nodes[it.ref];
        for (int c = 0; c < 4; ++c) {
            if (w.child[c] == kEmptyChild) continue;
            Item ch{w.child[c], {w.lox[c], w.loy[c], w.loz[c]}, {w.hix[c], w.hiy[c], w.hiz[c]}};
            for (int a = 0; a < 3; ++a)
                if (ch.lo[a] < it.lo[a] || ch.hi[a] > it.hi[a]) return 5;
            stack.push_back(ch);
        }
    }
    for (size_t i = 0; i < n; ++i)
        if (seen[i] != 1) return 6;
    // Quantized device trees: both hold exactly the float tree's leaves, and
    // along every root-to-leaf path the decoded child boxes contain the leaf's
    // float box (so the slab test over decoded boxes never culls a leaf the
    // float boxes would visit).
    std::map<uint32_t, std::array<float, 6>> leaf_box;
    for (const WNode& w : nodes)
        for (int c = 0; c < 4; ++c)
            if (w.child[c] != kEmptyChild && ref_is_leaf(w.child[c]))
                leaf_box[w.child[c]] = {w.lox[c], w.loy[c], w.loz[c], w.hix[c], w.hiy[c], w.hiz[c]};
    if (const int r = check_quantized(s.arr.bvh.q4, leaf_box)) return 8 + r;
    if (const int r = check_quantized(s.arr.bvh.q8, leaf_box)) return 10 + r;
    return 0;
}

void* emu_env_create(const crg_env_desc* d) {
    try {
        auto e = std::make_unique<EmuEnv>();
        e->he = build_env(*d);
        DevEnv& v = e->dev;
        v = DevEnv{};
        v.width = e->he.width;
        v.height = e->he.height;
        v.rotation = e->he.rotation;
        v.intensity = e->he.intensity;
        std::memcpy(v.constant, e->he.constant, sizeof(v.constant));
        v.is_image = e->he.is_image;
        if (e->he.is_image) {
            const size_t n = size_t(v.width) * v.height;
            e->texels.resize(n);
            for (size_t i = 0; i < n; ++i)
                e->texels[i] = make_float4(e->he.texels[4 * i], e->he.texels[4 * i + 1], e->he.texels[4 * i + 2], 1);
            v.texels = e->texels.data();
            v.pdf = e->he.pdf.data();
            v.threshold = e->he.threshold.data();
            v.guide = e->he.guide.data();
            v.rowcos = e->he.rowcos.data();
        }
        return e.release();
    } catch (const std::exception& ex) {
        lastError = ex.what();
        return nullptr;
    }
}
void emu_env_destroy(void* e) { delete static_cast<EmuEnv*>(e); }

// Exhaustive check of the integer inversion table: for every 24-bit draw k the
// device texel must equal the reference's upper_bound over its binary64 CDF
// (reference src/render.cpp:73-75).  Returns the number of mismatching k.
uint64_t emu_env_verify_inversion(void* ep) {
    const EmuEnv& e = *static_cast<EmuEnv*>(ep);
    if (!e.he.is_image) return 0;
    const std::vector<double>& cdf = e.he.cdf;
    uint64_t bad = 0;
    for (uint32_t k = 0; k < (1u << 24); ++k) {
        const double pick = double(float(k) * (1.0f / 16777216.0f)) * e.he.weight_sum;
        const size_t ref = std::min(size_t(std::upper_bound(cdf.begin(), cdf.end(), pick) - cdf.begin()), cdf.size() - 1);
        if (env_invert(e.dev, k) != ref) ++bad;
    }
    return bad;
}
void emu_env_sample(void* e, uint64_t* rng, float* dir, float* pdf) {
    Rng r{*rng};
    const V3 d = env_sample(static_cast<EmuEnv*>(e)->dev, r, *pdf);
    dir[0] = d.x; dir[1] = d.y; dir[2] = d.z;
    *rng = r.state;
}
// The FP32 alpha-mask decision (mask_fp32) against the binary64 one, for each
// ray and each alpha-mask primitive the ray certainly hits in FP32: counts[0]
// FP32 accepts, [1] FP32 rejects, [2] undecided, [3] FP32 decisions that
// contradict the binary64 test, [4] certain FP32 hits the binary64 test misses.
int emu_mask_check(void* sp, const crg_ray* rays, uint64_t n, uint64_t* counts) {
    const EmuScene& s = *static_cast<EmuScene*>(sp);
    for (uint64_t k = 0; k < n; ++k) {
        const crg_ray& r = rays[k];
        RayQuery q;
        setup_query(q, V3{r.origin[0], r.origin[1], r.origin[2]}, V3{r.direction[0], r.direction[1], r.direction[2]},
                    r.tmin, r.tmax, V3{r.origin_low[0], r.origin_low[1], r.origin_low[2]}, r.ignore);
        if (!q.valid) continue;
        for (uint32_t i = 0; i < s.dev.prim_count; ++i) {
            const PrimRec& p = s.dev.prims[i];
            if (!(p.flags & kPrimMaskTexture)) continue;
            const V3 va{p.ax, p.ay, p.az}, vb{p.bx, p.by, p.bz}, vc{p.cx, p.cy, p.cz};
            float lo = 0, hi = 0;
            if (fast_triangle(q, va, vb, vc, lo, hi) != kAccept) continue;
            const int mr = mask_fp32(s.dev, q, i, p.material, va, vb, vc);
            const ExactHit e = exact_triangle(q, va, vb, vc, sf::inf(0));
            if (!e.hit) {
                ++counts[4];
                continue;
            }
            bool bad = false;
            const bool keep = mask_accepts(s.dev, i, p.material, e.u, e.v, bad);
            ++counts[mr == kAccept ? 0 : mr == kReject ? 1 : 2];
            if (bad || (mr == kAccept && !keep) || (mr == kReject && keep)) ++counts[3];
        }
    }
    return 0;
}

// The device's binary64 normalization: float-float (dunit_ff) and the emulation (dunit), on
// vectors given as binary64 bit patterns (3 per vector).
void emu_dunit(const uint64_t* v, float* ff, float* chain, uint64_t n) {
    for (uint64_t i = 0; i < n; ++i) {
        const D3 d{sf::f64{v[3 * i]}, sf::f64{v[3 * i + 1]}, sf::f64{v[3 * i + 2]}};
        const V3 a = dunit_ff(d, V3{7, 7, 7}), b = dunit(d, V3{7, 7, 7});
        ff[3 * i] = a.x, ff[3 * i + 1] = a.y, ff[3 * i + 2] = a.z;
        chain[3 * i] = b.x, chain[3 * i + 1] = b.y, chain[3 * i + 2] = b.z;
    }
}

// surfaceAt's shading normal unit(n0 w + n1 hu + n2 hv): float-float (interp_unit_ff, else
// the chain) and the binary64 chain alone; per case 9 floats n0,n1,n2 and 2 binary64 weights
// hu, hv (w = (1 - hu) - hv in binary64, as the device forms it).
void emu_interp_unit(const float* n, const uint64_t* uv, float* ff, float* chain, uint64_t count) {
    for (uint64_t i = 0; i < count; ++i) {
        const V3 n0{n[9 * i], n[9 * i + 1], n[9 * i + 2]}, n1{n[9 * i + 3], n[9 * i + 4], n[9 * i + 5]},
            n2{n[9 * i + 6], n[9 * i + 7], n[9 * i + 8]};
        const sf::f64 hu{uv[2 * i]}, hv{uv[2 * i + 1]};
        const sf::f64 w = sf::sub(sf::sub(sf::one(), hu), hv);
        const D3 N = dadd(dadd(dscale(widen(n0), w), dscale(widen(n1), hu)), dscale(widen(n2), hv));
        const V3 fb{7, 7, 7};
        V3 a;
        if (!interp_unit_ff(n0, n1, n2, w, hu, hv, a)) a = dunit_ff(N, fb);
        const V3 b = dunit(N, fb);
        ff[3 * i] = a.x, ff[3 * i + 1] = a.y, ff[3 * i + 2] = a.z;
        chain[3 * i] = b.x, chain[3 * i + 1] = b.y, chain[3 * i + 2] = b.z;
    }
}

// Shading's surface reconstruction both ways for each ray's closest hit: from float-pair
// barycentrics with error bounds (hit_uv_ff + surface_at_ff, and hit_t_ff for float(t))
// and from the binary64 chain (exact_uv_of + surface_at).  Per ray 7 floats out each
// (position, shading normal, float(t)); decided bit 0: the pair path gave the surface,
// bit 1: it gave float(t); rays that miss get decided = 4.
int emu_hit_surface(void* sp, const crg_ray* rays, uint64_t n, float* ff, float* chain, uint8_t* decided) {
    const EmuScene& s = *static_cast<EmuScene*>(sp);
    for (uint64_t i = 0; i < n; ++i) {
        const crg_ray& r = rays[i];
        RayQuery q;
        setup_query(q, V3{r.origin[0], r.origin[1], r.origin[2]}, V3{r.direction[0], r.direction[1], r.direction[2]},
                    r.tmin, r.tmax, V3{r.origin_low[0], r.origin_low[1], r.origin_low[2]}, r.ignore);
        TraceCount tc{0, 0, 0};
        const ClosestResult b = traverse<false, true>(s.dev, q, tc);
        if (b.overflow || b.tex_error) {
            lastError = "traversal failed";
            return 1;
        }
        decided[i] = 4;
        if (b.prim == kNoTriangle) continue;
        const V3 d{r.direction[0], r.direction[1], r.direction[2]};
        const PrimRec& p = s.dev.prims[b.prim];
        HitUvFF h;
        Surf a{};
        float tf = 0.0f;
        const bool uv = hit_uv_ff(q, V3{p.ax, p.ay, p.az}, V3{p.bx, p.by, p.bz}, V3{p.cx, p.cy, p.cz}, h);
        const bool ok = uv && surface_at_ff(s.dev, b.prim, d, h, a);
        const bool ok_t = uv && hit_t_ff(q, h, tf);
        const ExactHit e = exact_uv_of(s.dev, q, b.prim, true);
        const Surf c = surface_at(s.dev, b.prim, d, e.u, e.v, false);
        const float ct = sf::to_float(e.t);
        if (!ok) a = c;
        if (!ok_t) tf = ct;
        decided[i] = (ok ? 1 : 0) | (ok_t ? 2 : 0);
        // The pairs' error bounds themselves: bit 3 when u, v, w or t lie farther
        // from the chain's binary64 values than claimed.
        if (uv) {
            auto dbl = [](sf::f64 x) { double d; std::memcpy(&d, &x.b, 8); return (long double)d; };
            auto pair = [](FF x) { return (long double)x.hi + (long double)x.lo; };
            const sf::f64 w = sf::sub(sf::sub(sf::one(), e.u), e.v);
            const long double du = fabsl(pair(h.u) - dbl(e.u)), dv = fabsl(pair(h.v) - dbl(e.v));
            const long double dw = fabsl((pair(h.w) - dbl(w)) + ((pair(h.u) - dbl(e.u)) + (pair(h.v) - dbl(e.v))));
            float et = 0.0f;
            const FF tp = hit_t_pair(q, h, et);
            bool bad = du > h.eu || dv > h.ev || dw > h.ew || fabsl(pair(tp) - dbl(e.t)) > et;
            if (bad) decided[i] |= 8;
        }
        const float fo[7] = {a.position.x, a.position.y, a.position.z, a.normal.x, a.normal.y, a.normal.z, tf};
        const float co[7] = {c.position.x, c.position.y, c.position.z, c.normal.x, c.normal.y, c.normal.z, ct};
        for (int k = 0; k < 7; ++k) {
            ff[7 * i + k] = fo[k];
            chain[7 * i + k] = co[k];
        }
    }
    return 0;
}

// surfaceAt's hit point float((a w + b hu) + c hv) per component: float-float
// (interp_point_ff, else the chain) and the binary64 chain alone; per case 9 floats a, b, c
// (triangle vertices) and 2 binary64 weights hu, hv (w = (1 - hu) - hv in binary64).
// decided[i] = 1 where the float-float path alone gave the result.
void emu_interp_point(const float* v, const uint64_t* uv, float* ff, float* chain, uint8_t* decided,
                      uint64_t count) {
    for (uint64_t i = 0; i < count; ++i) {
        const V3 a{v[9 * i], v[9 * i + 1], v[9 * i + 2]}, b{v[9 * i + 3], v[9 * i + 4], v[9 * i + 5]},
            c{v[9 * i + 6], v[9 * i + 7], v[9 * i + 8]};
        const sf::f64 hu{uv[2 * i]}, hv{uv[2 * i + 1]};
        const sf::f64 w = sf::sub(sf::sub(sf::one(), hu), hv);
        const V3 x = narrow(dadd(dadd(dscale(widen(a), w), dscale(widen(b), hu)), dscale(widen(c), hv)));
        FF fw, fu, fv;
        V3 f;
        const bool ok = ff_of(w, fw) && ff_of(hu, fu) && ff_of(hv, fv) && interp_point_ff(a, b, c, fw, fu, fv, f);
        if (!ok) f = x;
        decided[i] = ok ? 1 : 0;
        ff[3 * i] = f.x, ff[3 * i + 1] = f.y, ff[3 * i + 2] = f.z;
        chain[3 * i] = x.x, chain[3 * i + 1] = x.y, chain[3 * i + 2] = x.z;
    }
}

// One coordinate of offsetRay's shifted point: float-float (offset_coord) and the binary64
// emulation, for coordinate p, scale and normal component n.
void emu_offset(const float* p, const float* scale, const float* n, float* ff, float* chain, uint64_t count) {
    for (uint64_t i = 0; i < count; ++i) {
        const sf::f64 distance = sf::mul(sf::from_float(4 * kFltEps), sf::from_float(scale[i]));
        ff[i] = offset_coord(p[i], scale[i], n[i]);
        chain[i] = sf::to_float(sf::add(sf::from_float(p[i]), sf::mul(distance, sf::from_float(n[i]))));
    }
}

// The device's MIS weight: float-float (power_heuristic_ff) and the binary64 chain.
void emu_mis(const float* a, const float* b, float* ff, float* chain, uint64_t n) {
    for (uint64_t i = 0; i < n; ++i) {
        ff[i] = power_heuristic_ff(a[i], b[i]);
        chain[i] = power_heuristic(a[i], b[i]);
    }
}

// The device's x^5 for the Schlick term: float-float (pow5_ff) and the binary64 chain (pow5).
void emu_pow5(const float* x, float* ff, float* chain, uint64_t n) {
    for (uint64_t i = 0; i < n; ++i) {
        ff[i] = pow5_ff(x[i]);
        chain[i] = pow5(x[i]);
    }
}

// The device's expf (float-float with a binary64-emulated fallback, correctly rounded) over n arguments.
void emu_expf(const float* x, float* out, uint64_t n) {
    for (uint64_t i = 0; i < n; ++i) out[i] = expf_ff(x[i]);
}

// Slab test of one box, quantized by the builder as child 0 of a 4-wide node
// (with a sibling box as child 1 when sib_lo is not null, which moves the
// node's origin and coarsens its scale): 1 when node_hit reports child 0 as
// hit within [tmin, tmax].
int emu_box_hit(const float* o, const float* d, float tmin, float tmax, const float* lo, const float* hi,
                const float* sib_lo, const float* sib_hi) {
    RayQuery q;
    setup_query(q, V3{o[0], o[1], o[2]}, V3{d[0], d[1], d[2]}, tmin, tmax, V3{0, 0, 0}, kNoTriangle);
    const uint32_t refs[2] = {1, 2};
    float l[6] = {lo[0], lo[1], lo[2], 0, 0, 0}, h[6] = {hi[0], hi[1], hi[2], 0, 0, 0};
    int count = 1;
    if (sib_lo) {
        for (int a = 0; a < 3; ++a) {
            l[3 + a] = sib_lo[a];
            h[3 + a] = sib_hi[a];
        }
        count = 2;
    }
    const QNode4 n = quantize_boxes4(l, h, refs, count);
    float nears[4];
    return int(node_hit<4>(q, n, tmax, nears) & 1u);
}

// Film encoders on the CPU with the device's shared code (film.h) and the host
// half (film_host.cpp): display bytes by threshold search, RGBE and the
// scanline run-length code, the same checks and messages.
int emu_film_png8(const float* film, int width, int height, float exposure, uint8_t* rgb8) {
    try {
        check_dimensions(width, height, "PNG output");
        const size_t n = size_t(width) * size_t(height);
        for (size_t i = 0; i < 3 * n; ++i)
            if (!std::isfinite(film[i]) || film[i] < 0)
                throw std::runtime_error("PNG output: pixels must contain finite, nonnegative linear radiance");
        check_exposure(exposure);
        float t[255];
        display_thresholds(std::exp2(double(exposure)), t);
        for (size_t i = 0; i < 3 * n; ++i) rgb8[i] = uint8_t(display_byte(film[i], t));
        return 0;
    } catch (const std::exception& e) {
        lastError = e.what();
        return 1;
    }
}

int emu_film_hdr(const float* film, int width, int height, uint8_t* out, uint64_t capacity, uint64_t* size) {
    try {
        check_dimensions(width, height, "HDR output");
        const uint32_t w = uint32_t(width), h = uint32_t(height);
        for (size_t i = 0; i < size_t(w) * h; ++i) {
            const float* p = film + 3 * i;
            if (!std::isfinite(p[0]) || !std::isfinite(p[1]) || !std::isfinite(p[2]) || p[0] < 0 || p[1] < 0 || p[2] < 0)
                throw std::runtime_error("HDR output: pixels must contain finite, nonnegative linear radiance");
            if (std::max({p[0], p[1], p[2]}) >= std::ldexp(1.0f, 127))
                throw std::runtime_error("HDR output: radiance exceeds the Radiance RGBE exponent range");
        }
        const std::string head = hdr_header(width, height);
        std::vector<uint8_t> file(head.begin(), head.end());
        std::vector<uint8_t> planes(size_t(w) * 4), code(rle_bound(w));
        for (uint32_t y = 0; y < h; ++y) {
            for (uint32_t x = 0; x < w; ++x) {
                uint8_t q[4];
                const float* p = film + 3 * (size_t(y) * w + x);
                rgbe_of(p[0], p[1], p[2], q);
                if (w < 8 || w >= 32768) file.insert(file.end(), q, q + 4);
                for (int c = 0; c < 4; ++c) planes[size_t(c) * w + x] = q[c];
            }
            if (w < 8 || w >= 32768) continue;
            const uint8_t sh[4] = {2, 2, uint8_t((w >> 8) & 0xff), uint8_t(w & 0xff)};
            file.insert(file.end(), sh, sh + 4);
            for (int c = 0; c < 4; ++c) {
                const uint32_t k = rle_component(planes.data() + size_t(c) * w, 1u, w, code.data());
                file.insert(file.end(), code.begin(), code.begin() + k);
            }
        }
        *size = file.size();
        if (file.size() > capacity) throw std::runtime_error("HDR output buffer too small");
        std::memcpy(out, file.data(), file.size());
        return 0;
    } catch (const std::exception& e) {
        lastError = e.what();
        return 1;
    }
}

float emu_env_pdf(void* e, const float* d) { return env_pdf(static_cast<EmuEnv*>(e)->dev, V3{d[0], d[1], d[2]}); }
void emu_env_radiance(void* e, const float* d, float* rgb) {
    const V3 r = env_radiance(static_cast<EmuEnv*>(e)->dev, V3{d[0], d[1], d[2]});
    rgb[0] = r.x; rgb[1] = r.y; rgb[2] = r.z;
}

int emu_trace(void* sp, const crg_ray* rays, uint64_t n, crg_hit* hits, int any, uint64_t* fallbacks,
              uint64_t schedule) {
    const EmuScene& s = *static_cast<EmuScene*>(sp);
    uint64_t fbt = 0;
    for (uint64_t i = 0; i < n; ++i) {
        const crg_ray& r = rays[i];
        RayQuery q;
        setup_query(q, V3{r.origin[0], r.origin[1], r.origin[2]}, V3{r.direction[0], r.direction[1], r.direction[2]},
                    r.tmin, r.tmax, V3{r.origin_low[0], r.origin_low[1], r.origin_low[2]}, r.ignore);
        TraceCount tc{0, 0, 0};
        // FP32-only pass first, exact rerun only when flagged (the kernels' order).
        // Schedule bit 3 walks the 8-wide tree instead of the 4-wide one.
        uint64_t rng = schedule * 0x9e3779b97f4a7c15ull + i;
        ClosestResult b;
        Pending cands;
        const uint64_t sched = schedule & ~uint64_t(8);
        if (schedule & 8) {
            if (sched & 4)
                b = any ? phased<true, false, false, 8>(s.dev, q, tc, rng, &cands)
                        : phased<false, false, false, 8>(s.dev, q, tc, rng, &cands);
            else
                b = any ? phased<true, false, true, 8>(s.dev, q, tc, rng, &cands) : phased<false, false, true, 8>(s.dev, q, tc, rng, &cands);
        } else if (!sched) b = any ? traverse<true, false>(s.dev, q, tc, &cands) : traverse<false, false>(s.dev, q, tc, &cands);
        else if (sched & 4)  // the traversal variant without FP32 alpha-mask tests
            b = any ? phased<true, false, false>(s.dev, q, tc, rng, &cands) : phased<false, false, false>(s.dev, q, tc, rng, &cands);
        else if (sched & 1) b = any ? phased<true, false>(s.dev, q, tc, rng, &cands) : phased<false, false>(s.dev, q, tc, rng, &cands);
        else b = any ? phased<true, true>(s.dev, q, tc, rng, &cands) : phased<false, true>(s.dev, q, tc, rng, &cands);
        const int work = any ? exact_work<true>(b) : exact_work<false>(b);
        const bool rerun = work != 0;
        if (rerun) {
            const Pending pd = pending_of(b, work);
            setup_query(q, V3{r.origin[0], r.origin[1], r.origin[2]}, V3{r.direction[0], r.direction[1], r.direction[2]},
                        r.tmin, r.tmax, V3{r.origin_low[0], r.origin_low[1], r.origin_low[2]}, r.ignore);
            b = any ? exact_pass<true>(s.dev, q, pd, tc) : exact_pass<false>(s.dev, q, pd, tc);
        }
        fbt += rerun ? 1 : 0;
        if (b.overflow) {
            lastError = "BVH traversal exceeded its bounded stack";
            return 1;
        }
        if (b.tex_error) {
            lastError = "Non-finite texture coordinates or transform";
            return 1;
        }
        crg_hit& h = hits[i];
        h.hit = b.prim != kNoTriangle;
        h.triangle = h.hit ? b.id : kNoTriangle;
        h.t = h.u = h.v = 0;
        if (h.hit) {
            if (!b.exact) {
                const ExactHit e = exact_of(s.dev, q, b.prim);
                b.t = e.t;
                b.u = e.u;
                b.v = e.v;
            }
            std::memcpy(&h.t, &b.t.b, 8);
            std::memcpy(&h.u, &b.u.b, 8);
            std::memcpy(&h.v, &b.v.b, 8);
        }
    }
    if (fallbacks) *fallbacks = fbt;
    return 0;
}

// Device camera / offset / BSDF pieces for golden comparisons.
int emu_camera_generate(const crg_camera* c, float px, float py, int w, int h, crg_ray* out) {
    try {
        const DevCamera cam = prepare_camera(*c, w, h);
        V3 o, d;
        float tmin, tmax;
        if (!camera_ray(cam, px, py, o, d, tmin, tmax)) throw std::invalid_argument("Unrepresentable perspective ray");
        std::memset(out, 0, sizeof(*out));
        out->origin[0] = o.x; out->origin[1] = o.y; out->origin[2] = o.z;
        out->direction[0] = d.x; out->direction[1] = d.y; out->direction[2] = d.z;
        out->tmin = tmin;
        out->tmax = tmax;
        out->ignore = kNoTriangle;
        return 0;
    } catch (const std::exception& e) {
        lastError = e.what();
        return 1;
    }
}

// Camera::generate on the device for many pixels: the float-float path with the chain for hard
// cases (camera_ray) and the binary64 chain alone; 6 floats per pixel (direction, tmin, tmax,
// ok). Returns 1 when the camera itself is invalid.
int emu_camera_pair(const crg_camera* c, const float* px, const float* py, uint64_t n, int w, int h, float* ff,
                    float* chain) {
    try {
        const DevCamera cam = prepare_camera(*c, w, h);
        for (uint64_t i = 0; i < n; ++i) {
            V3 o, d{0, 0, 0};
            float t0 = 0, t1 = 0;
            bool ok = camera_ray(cam, px[i], py[i], o, d, t0, t1);
            float* a = ff + 6 * i;
            a[0] = d.x, a[1] = d.y, a[2] = d.z, a[3] = ok ? t0 : 0, a[4] = ok ? t1 : 0, a[5] = ok;
            d = V3{0, 0, 0}, t0 = t1 = 0;
            ok = camera_ray_chain(cam, px[i], py[i], o, d, t0, t1);
            float* b = chain + 6 * i;
            b[0] = d.x, b[1] = d.y, b[2] = d.z, b[3] = ok ? t0 : 0, b[4] = ok ? t1 : 0, b[5] = ok;
        }
        return 0;
    } catch (const std::exception& e) {
        lastError = e.what();
        return 1;
    }
}

int emu_render(void* sp, void* ep, const crg_camera* camera, const crg_options* o, float* film, crg_stats* stats) {
    const EmuScene& s = *static_cast<EmuScene*>(sp);
    const EmuEnv& env = *static_cast<EmuEnv*>(ep);
    try {
        validate_options(*o);
        DevJob J{};
        J.scene = s.dev;
        J.env = env.dev;
        const LightTable lt = build_light_table(s.lights, s.arr.tri_to_prim, env.he.integral);
        J.lights = DevLights{lt.entries.data(), uint32_t(lt.entries.size()), lt.env_probability};
        J.cam = prepare_camera(*camera, o->width, o->height);
        J.width = o->width;
        J.height = o->height;
        J.spp = o->samples;
        J.bounces = o->bounces;
        J.seed = o->seed;
        J.nee = o->next_event_estimation;
        J.rr = o->russian_roulette;
        const uint32_t npix = uint32_t(o->width) * uint32_t(o->height);
        std::atomic<uint32_t> next{0};
        std::atomic<uint64_t> replays{0}, ext{0}, shd{0};
        std::atomic<uint64_t> best_key{~0ull};
        std::string err_text;
        std::mutex mu;
        auto work = [&] {
            for (;;) {
                const uint32_t p = next.fetch_add(1);
                if (p >= npix) break;
                sf::f64 acc[3] = {sf::zero(0), sf::zero(0), sf::zero(0)};
                for (int smp = 0; smp < o->samples; ++smp) {
                    PathState st;
                    bool ok = begin_sample(J, p, uint32_t(smp), false, st);
                    ErrorOut err{};
                    uint32_t code = ok ? 0 : uint32_t(kErrCamera);
                    while (ok) {
                        RayQuery q;
                        setup_query(q, st.o, st.d, st.tmin, st.tmax, st.lo, st.prev);
                        TraceCount tc{0, 0, 0};
                        // Same two-phase order as k_extend / k_extend_exact.
                        Pending cands;
                        ClosestResult c = traverse<false, false>(s.dev, q, tc, &cands);
                        if (const int work = exact_work<false>(c)) {
                            RayQuery qx;
                            setup_query(qx, st.o, st.d, st.tmin, st.tmax, st.lo, st.prev);
                            c = exact_pass<false>(s.dev, qx, pending_of(c, work), tc);
                        }
                        ++ext;
                        if (c.overflow || c.tex_error) {
                            code = c.overflow ? kErrStack : kErrTexture;
                            break;
                        }
                        ShadowOut sh;
                        const uint32_t a = bounce(J, st, c.prim, sh, err);
                        if (err.code) {
                            code = err.code;
                            break;
                        }
                        if (sh.emit) {
                            RayQuery sq;
                            setup_query(sq, sh.o, sh.d, 0, sh.tmax, sh.lo, sh.ignore);
                            ClosestResult b = traverse<true, false>(s.dev, sq, tc, &cands);
                            if (const int work = b.overflow ? 0 : exact_work<true>(b)) {
                                RayQuery sx;
                                setup_query(sx, sh.o, sh.d, 0, sh.tmax, sh.lo, sh.ignore);
                                b = exact_pass<true>(s.dev, sx, pending_of(b, work), tc);
                            }
                            ++shd;
                            if (b.tex_error) {
                                code = kErrTexture;
                                break;
                            }
                            if (b.prim == kNoTriangle) st.L = st.L + sh.c;
                        }
                        if (a == kReplay) {
                            ++replays;
                            begin_sample(J, p, uint32_t(smp), true, st);
                            continue;
                        }
                        if (a == kRetire) break;
                    }
                    if (!code && !radiance_ok(st.L)) code = kErrBadRadiance;
                    if (code) {
                        const uint64_t key = uint64_t(p) * uint64_t(o->samples) + uint64_t(smp);
                        std::lock_guard<std::mutex> lk(mu);
                        if (key < best_key.load()) {
                            best_key = key;
                            ErrorRec rec{};
                            rec.key = key;
                            rec.code = code;
                            rec.depth = err.depth;
                            rec.triangle = err.triangle;
                            rec.material = err.material;
                            rec.tbits = err.t.b;
                            const V3 vs[3] = {err.point, err.origin, err.direction};
                            float* dst[3] = {rec.point, rec.origin, rec.direction};
                            for (int k = 0; k < 3; ++k) {
                                dst[k][0] = vs[k].x;
                                dst[k][1] = vs[k].y;
                                dst[k][2] = vs[k].z;
                            }
                            err_text = format_path_error(rec, o->width, o->samples);
                        }
                    }
                    acc[0] = sf::add(acc[0], sf::from_float(st.L.x));
                    acc[1] = sf::add(acc[1], sf::from_float(st.L.y));
                    acc[2] = sf::add(acc[2], sf::from_float(st.L.z));
                }
                const sf::f64 n = sf::from_i32(o->samples);
                for (int k = 0; k < 3; ++k) film[3 * size_t(p) + k] = sf::to_float(sf::div(acc[k], n));
            }
        };
        const int workers = int(std::max(1u, std::thread::hardware_concurrency()));
        std::vector<std::thread> pool;
        for (int i = 0; i < workers; ++i) pool.emplace_back(work);
        for (auto& t : pool) t.join();
        if (!err_text.empty()) throw std::runtime_error(err_text);
        if (stats) {
            std::memset(stats, 0, sizeof(*stats));
            stats->camera_samples = uint64_t(npix) * o->samples;
            stats->boundary_retries = replays.load();
            stats->extend_rays = ext.load();
            stats->shadow_rays = shd.load();
        }
        return 0;
    } catch (const std::exception& e) {
        lastError = e.what();
        return 1;
    }
}
}
