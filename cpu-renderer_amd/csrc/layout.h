// layout.h — device-resident data layout of the wavefront path tracer.
//
// Scene (per resident scene, all in HBM):
//   WNode   128-byte 4-wide BVH node: child boxes as SoA lanes + child refs.
//   PrimRec 48-byte traversal record per triangle in leaf order (v0,id | v1,flags | v2,material).
//   PrimAux 64-byte shading record per triangle in leaf order (vertex normals,
//           geometric normal, unit geometric normal, area); fetched only by shade.
// Environment: texels + per-texel float pdf + integer inversion thresholds.
// Path state: structure-of-arrays, one slot per in-flight camera sample.
#pragma once
#include <stdint.h>

#include "vmath.h"

#if defined(__HIPCC__) || defined(__HIP__)
#include <hip/hip_runtime.h>
#else
// Plain C++ builds of the device logic (CPU emulation in tests) use this
// stand-in for HIP's vector type.
struct alignas(16) float4 {
    float x, y, z, w;
};
inline float4 make_float4(float x, float y, float z, float w) { return float4{x, y, z, w}; }
#endif

namespace cr {

// Scene records are read through references in the global address space.
// Scene pointers are themselves read from memory, so the compiler cannot prove
// that they point to global memory and would emit flat loads, which also count
// against the LDS/scalar-load counter and so stall LDS traffic behind them.
// Members are read through the reference (a copy of the whole record would go
// through a generic reference again).
#if defined(__HIP_DEVICE_COMPILE__)
#define CR_GLOBAL __attribute__((address_space(1)))
#else
#define CR_GLOBAL
#endif
template <class T>
CR_HDI const CR_GLOBAL T& gref(const T* p) {
    return *(const CR_GLOBAL T*)p;
}

// A query the FP32-only traversal could not decide alone: its certain nearest
// hit so far and up to kCand candidate primitives (undecided tests, alpha-mask
// primitives, certain hits whose order against the best was undecided), for
// the exact pass to resolve in binary64 without traversing again; `full` set
// when more than kCand candidates were met (exact traversal instead).
#ifndef CR_CAND
#define CR_CAND 4
#endif
constexpr uint32_t kCand = CR_CAND;
struct alignas(16) Pending {
    uint32_t best;  // leaf-order primitive, kNoTriangle when none
    uint32_t n;     // candidates stored
    float thi;      // upper bound of the best's t (the query's tmax when none)
    uint32_t full;
    uint32_t cand[kCand];
    float lo[kCand];  // lower bound of each candidate's t (0 when unknown)
};

// (cos, sin) of the sampling angle phi = 2*pi*u for one 24-bit draw u = k/2^24;
// the table holds all kTurnEntries draws, evaluated once on the host.
struct alignas(8) TurnCS {
    float c, s;
};
constexpr uint32_t kTurnEntries = 1u << 24;

// prim_ng entries read as four plain floats (geometric normal, unused w).
struct PrimAuxNg {
    float x, y, z, w;
};

// Child reference: bit 31 clear -> inner node index; bit 31 set -> leaf with
// count-1 in bits 27..30 (1..16 primitives) and the first primitive in bits 0..26.
constexpr uint32_t kEmptyChild = 0x7fffffffu;
constexpr uint32_t kLeafBit = 0x80000000u;
constexpr uint32_t kLeafMaxCount = 16;
constexpr uint32_t kLeafMaxFirst = (1u << 27) - 1;
CR_HDI bool ref_is_leaf(uint32_t r) { return (r & kLeafBit) != 0; }
CR_HDI uint32_t leaf_first(uint32_t r) { return r & kLeafMaxFirst; }
CR_HDI uint32_t leaf_count(uint32_t r) { return ((r >> 27) & 15u) + 1; }
CR_HDI uint32_t make_leaf(uint32_t first, uint32_t count) { return kLeafBit | ((count - 1) << 27) | first; }

// 4-wide node with float child boxes: the host builder's tree, from which the
// device's quantized trees are made.
struct alignas(128) WNode {
    float lox[4], hix[4], loy[4], hiy[4], loz[4], hiz[4];
    uint32_t child[4];
    uint32_t pad[4];
};
static_assert(sizeof(WNode) == 128, "WNode must be one 128-byte line");

// Device node of a W-wide tree (W = 4: 64 bytes, W = 8: 128 bytes) with child
// boxes quantized to 8 bits per plane.  Per axis a the node stores an origin
// p_a (the node box's low corner) and a power-of-two scale s_a = 2^(e_a - 127)
// (the float whose exponent field is e_a, 1 <= e_a <= 254); a child's planes
// decode as
//     lo = p_a + float(qlo) * s_a,   hi = p_a + float(qhi) * s_a
// where the product is exact (q < 2^8, s_a a power of two) and the sum rounds
// once in binary32.  The builder evaluates exactly this expression and picks
// qlo / qhi so that the decoded box contains the child's float box, so the
// slab test on decoded boxes over-includes as the float boxes did.  Byte k of
// word qlo[a][c / 4] is child c's plane (c % 4 == k).  Empty children have
// child ref kEmptyChild.
template <int W>
struct alignas(16 * W) QNode {
    float px, py, pz;
    uint32_t exps;  // e_x | e_y << 8 | e_z << 16 | child count << 24
    uint32_t qlo[3][W / 4];
    uint32_t qhi[3][W / 4];
    uint32_t child[W];
};
using QNode4 = QNode<4>;
using QNode8 = QNode<8>;
static_assert(sizeof(QNode4) == 64, "4-wide node is one 64-byte block");
static_assert(sizeof(QNode8) == 128, "8-wide node is one 128-byte line");

CR_HDI float qscale(uint32_t e) { return sf::bitsf(e << 23); }
// A child plane: p + q * 2^(e-127).  The product of an 8-bit q and a power of
// two is exact, so the fused form rounds once, like the plain sum (one
// instruction on the device); host and device decode with this same function.
CR_HDI float qdecode(float p, uint32_t q, float s) { return __builtin_fmaf(float(q), s, p); }

enum PrimFlags : uint32_t {
    kPrimCull = 1u,        // single-sided opaque: reject when dot(Ng, d) >= 0
    kPrimNeverHit = 2u,    // MASK material whose constant alpha is below its cutoff
    kPrimMaskTexture = 4u,  // MASK material with a base-colour texture: alpha sampled at the exact hit
    kPrimClassShift = 8u,   // bits 8..9: shading class (scheduling only, never results)
};

// Shading classes: shade groups the paths of a chunk by class so that a wave
// shades one kind of surface (or only misses) at a time.
enum ShadeClass : uint32_t {
    kShadeMiss = 0,        // environment / no hit / failed query
    kShadeOpaque = 1,      // opaque, non-metal
    kShadeMetal = 2,       // opaque, metallic > 0
    kShadeTransmissive = 3,
    kShadeClasses = 4,
};

struct alignas(16) PrimRec {
    float ax, ay, az;
    uint32_t id;  // triangle index in the finalized scene (tie-break + output)
    float bx, by, bz;
    uint32_t flags;
    float cx, cy, cz;
    uint32_t material;
};
static_assert(sizeof(PrimRec) == 48, "PrimRec is three 16-byte vectors");

struct alignas(16) PrimAux {
    float n0[3], n1[3], n2[3];
    float ng[3];   // geometric normal as stored by finalize
    float ung[3];  // unit(ng) evaluated in binary64 (offsetRay's normalization)
    float area;
};
static_assert(sizeof(PrimAux) == 64, "PrimAux is one 64-byte record");

// Texture binding (reference TextureSlot, src/scene.h:50-53) with the
// binary64 cos/sin of its rotation computed on the host (std::cos/std::sin on
// the same value the reference converts).
struct DevSlot {
    int32_t texture;   // -1: no map
    int32_t texcoord;  // UV set 0 or 1
    float offset[2], scale[2];
    uint64_t cos_bits, sin_bits;
};

struct DevImage {
    uint64_t offset;       // first texel in the scene's texel array
    uint64_t srgb_offset;  // first texel of the sRGB-decoded copy (images sampled by colour slots)
    int32_t width, height;
};

struct DevTexture {
    int32_t image, wrap_s, wrap_t, linear;  // wrap: 0 repeat, 1 clamp, 2 mirror
};

struct DevMaterial {
    float base[4];
    float metallic, roughness;
    float emission[3];
    float transmission, ior, thickness, att_distance;
    float att_color[3];
    float sigma[3];  // absorptionCoefficient(att_color, att_distance), evaluated on the host
    uint32_t double_sided;
    uint32_t maps;   // any slot bound (surfaceAt samples textures)
    float normal_scale;
    float alpha_cutoff;  // MASK materials
    DevSlot base_tex, mr_tex, normal_tex, emissive_tex, transmission_tex;
};

struct DevLightEntry {
    uint32_t type;       // 0 punctual, 1 triangle, 2 environment
    uint32_t index;      // punctual index or prim index
    uint32_t threshold;  // entry chosen when 24-bit draw k <= threshold
    float probability;
};

struct DevPunctual {
    uint32_t type;
    float position[3], direction[3], color[3];
    float intensity, range, cos_outer, cos_inner;
};

struct DevScene {
    const QNode4* nodes4;  // the 4-wide tree
    const QNode8* nodes8;  // the 8-wide tree over the same leaves (a collapse of the 4-wide one)
    const PrimRec* prims;
    const float4* prim_ng;        // geometric normal per prim (cull test)
    const PrimAux* aux;
    const DevMaterial* materials;
    const uint32_t* light_rank;   // per prim: entry index in the light table, or kNoTriangle
    const DevPunctual* punctual;
    const float4* uvs;       // per prim 3 x (uv0.xy, uv1.xy) when a material samples a texture
    const float4* tangents;  // per prim 3 x (tangent.xyz, handedness) when a material has a normal map
    const float4* texels;    // all images, row-major
    const DevImage* images;
    const DevTexture* textures;
    uint32_t node_count4, prim_count, punctual_count, node_count8;
};

// Light-selection mixture of one (scene, environment) pair.
struct DevLights {
    const DevLightEntry* entries;
    uint32_t count;
    float env_probability;
};

// Draws per guide bucket of the environment inversion: 2^24 >> kGuideShift
// buckets, each holding the texel index at its first draw, so that a draw's
// texel lies between two neighbouring entries (a few texels apart even in the
// sky's low-weight regions: one or two threshold probes instead of a dozen).
constexpr uint32_t kGuideShift = 2;
constexpr uint32_t kGuideEntries = (1u << (24 - kGuideShift)) + 1;

struct DevEnv {
    const float4* texels;       // rgb, 1
    const float* pdf;           // per-texel solid-angle pdf (binary64 on host, rounded)
    const uint32_t* threshold;  // T_i: first 24-bit draw whose CDF pick reaches texel i
    const uint32_t* guide;      // kGuideEntries entries: texel index at k = j << kGuideShift
    const float* rowcos;        // cosf(pi*y/H), y = 0..H
    int32_t width, height;
    float rotation, intensity;
    float constant[3];
    int32_t is_image;
};

struct DevCamera {
    float position[3], f[3], r[3], u[3];
    sf::f64 ratio, half_height;
    float near_clip, far_clip;
    int32_t width, height;
};

// One render request: scene/environment/camera/options plus its sample buffer.
struct DevJob {
    // Copies of the resident scene's and environment's descriptors: a path
    // reaches their arrays with the job record's load instead of a second,
    // dependent load of a descriptor it points to.
    DevScene scene;
    DevEnv env;
    DevLights lights;
    DevCamera cam;
    int32_t width, height, spp, bounces;
    uint64_t seed;
    int32_t nee, rr;
    float4* samples;     // [sample][window pixel] radiance (xyz)
    uint64_t item_base;  // first global work item of this job
    uint64_t items;      // win_w*win_h*spp
    // Window of the frame this job renders (the whole frame: 0, 0, width,
    // height); pixels keep their frame index, so their sample streams are the
    // full frame's.
    int32_t win_x, win_y, win_w, win_h;
};

enum PathFlag : uint32_t {
    kDepthMask = 0xffu,
    kLastDelta = 1u << 8,
    kPrecise = 1u << 9,
    kFailed = 1u << 10,  // camera ray failed: retire with zero radiance
};

// Per-path state as two aligned records.  Kernels reach paths through
// shuffled queues, so a field-per-array layout costs one cache line per field
// per path; a record costs one line.  The ray record is all the traversal
// kernels read.
struct alignas(64) PathRay {
    float4 o;        // origin, tmin
    float4 d;        // direction, tmax
    float4 lo;       // originLow, (unused)
    uint32_t prev;   // departed triangle id (precise replay) or kNoTriangle
    uint32_t job;    // job-table slot
    uint32_t hit;    // encoded closest hit of the last extend
    uint32_t flags;  // depth | kLastDelta | kPrecise | kFailed
};
static_assert(sizeof(PathRay) == 64, "PathRay is one 64-byte record");

// The first 64 bytes hold what the retire kernel reads (radiance, pixel and
// sample), so it fetches one 64-byte sector of the record.
struct alignas(128) PathRec {
    float4 thr;      // throughput, etaScale
    float4 rad;      // radiance, lastPdf
    uint64_t rng;
    uint32_t pixel, sample;
    int32_t medium;
    uint32_t pad0[3];
    float4 lastpos;  // lastPosition
    float4 absorb;   // active medium absorption
    uint32_t pad[8];
};
static_assert(sizeof(PathRec) == 128, "PathRec is one 128-byte record");

struct PathSoA {
    PathRay* ray;
    PathRec* rec;
};

struct ShadowSoA {
    uint32_t* path;
    uint32_t* job;  // job of the emitting path (read by the traversal without a gather)
    float4* o;   // origin, tmin
    float4* d;   // direction, tmax
    float4* lo;  // originLow, ignore (as bits)
    float4* c;   // contribution
};

struct ErrorRec {
    uint64_t key;    // pixel*spp + sample (smallest reported)
    uint64_t tbits;  // binary64 hit distance
    uint32_t job, code, depth, triangle, material, pad;
    float point[3], origin[3], direction[3], pad2;
};

enum ErrorCode : uint32_t {
    kErrNonFiniteThroughput = 1,
    kErrBadRadiance = 2,
    kErrNested = 3,
    kErrExitWithoutEntry = 4,
    kErrInvalidSecondary = 5,
    kErrOffsetOverflow = 6,
    kErrCamera = 7,
    kErrStack = 8,
    kErrInternal = 9,  // work item outside the job it maps to (engine bookkeeping)
    kErrTexture = 10,  // non-finite texture coordinates
};

constexpr int kMaxErrors = 64;  // error ring (records of one host step)

// Device job table: job sequence number s lives in slot s % kJobSlots; the
// live sequence range [seq_lo, seq_hi) is ordered by item_base.
constexpr uint32_t kJobSlots = 1024;

// Wavefront control block.  Queue counters are per iteration; issued/total
// count global work items (camera samples) over every job ever admitted.
// The traversal queues are split into this many contiguous parts, claimed
// through separate counters (workgroups spread over the XCDs round-robin, so
// each XCD starts on a part of its own).
#ifndef CR_CLAIM_PARTS
#define CR_CLAIM_PARTS 8
#endif
constexpr uint32_t kClaimParts = CR_CLAIM_PARTS;

struct Control {
    uint32_t qcount[2];
    uint32_t scount, rcount;
    uint32_t xcount, xscount;  // extend / shadow rays needing the exact traversal
    uint32_t fcount;           // idle path slots on the free stack
    uint32_t refill_n;         // free slots the next refill starts
    uint32_t refill_from;      // free-stack index of the first of them
    uint32_t refill_q;         // next-queue index of the first of them
    uint32_t seq_lo, seq_hi;   // live job sequence range
    uint32_t err_count;        // monotonic; records go to errors[n % kMaxErrors]
    uint32_t ehead, shead;     // (unused; the traversal queues are claimed through eparts / sparts)
    uint32_t khead;            // next unclaimed shading chunk (extend-queue index)
    uint32_t pool;             // path slots (capacity of every queue)
    uint32_t masks;            // a live job's scene has alpha-mask primitives: traversal variant with FP32 mask tests
    unsigned long long issued;
    unsigned long long total;
    unsigned long long refill_base;  // first item of the refill range
    unsigned long long fallbacks;    // exact triangle tests of k_trace_rays
    ErrorRec errors[kMaxErrors];
    // Claim heads of the extend / shadow queues, one per part, each on a
    // 128-byte line of its own (see the traversal kernels' claim).
    uint32_t eparts[kClaimParts * 32];
    uint32_t sparts[kClaimParts * 32];
};

// Per-job counters, one array per quantity, indexed by job slot.
struct JobCounters {
    unsigned long long retired[kJobSlots];
    unsigned long long replays[kJobSlots];
    unsigned long long extend_rays[kJobSlots];
    unsigned long long shadow_rays[kJobSlots];
    unsigned long long extend_nodes[kJobSlots];
    unsigned long long extend_tris[kJobSlots];
    unsigned long long shadow_nodes[kJobSlots];
    unsigned long long shadow_tris[kJobSlots];
    unsigned long long fallbacks[kJobSlots];
    unsigned long long exact_rays[kJobSlots];
};

}  // namespace cr
