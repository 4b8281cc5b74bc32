// kernels.h — host-side launch interface of the gfx950 kernels.
#pragma once
#include <hip/hip_runtime.h>

#include "layout.h"
#include "../../include/crgpu.h"

namespace cr {

// Byte-compatible with crg_ray / crg_hit of include/crgpu.h.
struct RayIn {
    float o[3], d[3];
    float tmin, tmax;
    float lo[3];
    uint32_t ignore;
};
struct HitOut {
    uint32_t hit, triangle;
    uint64_t t, u, v;  // binary64 bit patterns
};
static_assert(sizeof(RayIn) == 48, "crg_ray layout");
static_assert(sizeof(HitOut) == 32, "crg_hit layout");

// Bvh::intersect per ray over the scene's `width`-wide tree (4 or 8).
hipError_t launch_trace_rays(const DevScene* scene, const RayIn* rays, uint64_t n, HitOut* hits, int any,
                             Control* ctl, int width, hipStream_t st);

// Fixed per-device scratch of the allocation-free ray query: the control block
// (rounded up to 256 bytes), then the rays, then the hits of at most
// kNoAllocMaxRays queries (5.3 MB in all).
constexpr uint64_t kNoAllocMaxRays = uint64_t(1) << 16;
constexpr uint64_t kNoAllocCtlBytes = (sizeof(Control) + 255) / 256 * 256;
constexpr uint64_t kNoAllocScratchBytes = kNoAllocCtlBytes + kNoAllocMaxRays * (sizeof(RayIn) + sizeof(HitOut));
// Device address of that scratch on `device`, which must be the current
// device on the first call for it (hipGetSymbolAddress once per device).
hipError_t noalloc_scratch(int device, void** out);
// Per-function probe (crg_device_probe): applies one device function of the
// path to n input records (kinds and record layouts as in crgpu.h).
struct ProbeArgs {
    int what;
    const DevScene* scene;
    const DevEnv* env;
    DevCamera camera;
    const void* in;
    void* out;
    uint64_t n;
    const uint32_t* tri_to_prim;  // surfaceAt: finalized triangle -> leaf-order primitive
    const float* ung;             // offsetRay: unit(geometric normal) per record, from the host
};
hipError_t launch_probe(const ProbeArgs& a, hipStream_t st);
// Zeroes n 32-bit words with a kernel (no memset node).
hipError_t launch_clear_words(uint32_t* p, uint32_t n, hipStream_t st);
// The kernels of one wavefront iteration, in launch order.
enum Stage {
    kStageExtend = 0,
    kStageExtendExact,
    kStageShade,
    kStageShadow,
    kStageShadowExact,
    kStageRetire,
    kStageReset,
    kStageRefill,
    kStageCount
};
constexpr const char* kStageNames[kStageCount] = {"k_extend",       "k_extend_exact", "k_shade",
                                                  "k_shadow",       "k_shadow_exact", "k_retire",
                                                  "k_reset",        "k_refill"};

// Path-id queues of the wavefront loop.
struct Queues {
    uint32_t* q[2];   // extend queues (current / next iteration)
    uint32_t* retire; // finished paths
    uint32_t* retire_job;  // job slot of each retire entry (written with it)
    uint32_t* xq;     // extend rays needing the exact pass
    uint32_t* xsq;    // shadow records needing the exact pass
    uint32_t* freeq;  // stack of idle path slots
    Pending* xpend;   // per xq entry: what the FP32 pass left undecided
    Pending* xspend;  // per xsq entry
};

// Launch shape of the wavefront kernels (persistent-style grids that loop
// over their queue).
struct LaunchShape {
    unsigned grid;   // workgroups
    unsigned block;  // threads per workgroup: 64, 128 or 256
    bool count;      // traversal kernels with per-query work counters
    int width;       // BVH width the traversal kernels walk: 4 or 8
};

hipError_t launch_stage(Stage stage, const PathSoA& P, const ShadowSoA& S, const DevJob* jobs, const Queues& Q,
                        int parity, Control* ctl, JobCounters* cnt, LaunchShape shape, hipStream_t st);
hipError_t launch_pool_init(uint32_t* freeq, uint32_t pool, Control* ctl, hipStream_t st);
hipError_t launch_admit(Control* ctl, JobCounters* cnt, uint64_t total, uint32_t seq_lo, uint32_t seq_hi, uint32_t slot,
                        hipStream_t st);
hipError_t launch_accumulate(const DevJob* jobs, uint32_t job, uint64_t npix, float* film, hipStream_t st);
// Selects the traversal kernels' variant: with FP32 alpha-mask tests (masks != 0) or without.
hipError_t launch_set_masks(Control* ctl, uint32_t masks, hipStream_t st);
// Film encode: display bytes (ldr, 3 per pixel; skipped when null) and RGBE in
// per-row planes (rgbe: row-major rows of 4 planes of `width` bytes; skipped
// when null); first_bad[0] / [1] (preset to ~0u) receive the lowest pixel index
// that is non-finite or negative / whose largest component reaches 2^127.
// thresholds: 255 floats on the device.
hipError_t launch_film_encode(const float* film, uint64_t npix, uint32_t width, const float* thresholds, uint8_t* ldr,
                              uint8_t* rgbe, uint32_t* first_bad, hipStream_t st);
// Run-length codes every (row, plane) of `rgbe` into streams[(row*4+plane)*rle_bound(width)]
// and its byte count into lens[row*4+plane].
hipError_t launch_film_rle(const uint8_t* rgbe, uint32_t width, uint32_t height, uint8_t* streams, uint32_t* lens,
                           hipStream_t st);
// Points the kernels of the current device at its sampling-angle table
// (kTurnEntries entries in device memory, filled by fill_turn_table).
hipError_t set_turn_table(const TurnCS* table);

}  // namespace cr
