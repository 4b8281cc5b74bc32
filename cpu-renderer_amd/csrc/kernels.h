// kernels.h — host-side launch interface of the gfx950 kernels.
#pragma once
#include <hip/hip_runtime.h>

#include "layout.h"

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

hipError_t launch_trace_rays(const DevScene* scene, const RayIn* rays, uint64_t n, HitOut* hits, int any,
                             Control* ctl, hipStream_t st);
hipError_t launch_seed(const PathSoA& P, const DevJob* jobs, uint32_t njobs, uint32_t pool, uint32_t* q0,
                       Control* ctl, hipStream_t st);
// The kernels of one wavefront iteration, in launch order.
enum Stage {
    kStageExtend = 0,
    kStageExtendExact,
    kStageShade,
    kStageShadow,
    kStageShadowExact,
    kStageRetire,
    kStageReset,
    kStageCount
};
constexpr const char* kStageNames[kStageCount] = {"k_extend", "k_extend_exact", "k_shade", "k_shadow",
                                                  "k_shadow_exact", "k_retire", "k_reset"};

// Path-id queues of the wavefront loop.
struct Queues {
    uint32_t* q[2];   // extend queues (current / next iteration)
    uint32_t* retire; // finished paths
    uint32_t* xq;     // extend rays needing the exact traversal
    uint32_t* xsq;    // shadow records needing the exact traversal
};

hipError_t launch_stage(Stage stage, const PathSoA& P, const ShadowSoA& S, const DevJob* jobs, uint32_t njobs,
                        const Queues& Q, int parity, Control* ctl, unsigned grid, hipStream_t st);
hipError_t launch_accumulate(const DevJob* jobs, uint32_t job, uint64_t npix, float* film, hipStream_t st);

}  // namespace cr
