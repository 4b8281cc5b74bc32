// kernels.hip — gfx950 wavefront path-tracing kernels.
//
// One render (or a batch of renders) is a pool of path slots processed in
// iterations; each iteration is
//   extend  : closest-hit traversal of every queued ray          (Bvh::intersect)
//   shade   : hit -> surface, emission/MIS, NEE light sample, BSDF sample,
//             media, Russian roulette; emits shadow rays and next rays
//   shadow  : any-hit visibility of the NEE rays, adds their contribution
//   retire  : finished paths write their radiance into the job's sample
//             buffer and the slot immediately takes the next work item
// All per-path state lives in HBM as structure-of-arrays, so the result of a
// path never depends on which slot, wave or iteration ran it; the film is
// summed per pixel in sample order afterwards (accumulate), exactly like the
// reference's per-pixel double accumulation (reference src/render.cpp:384-415).
#include <hip/hip_runtime.h>

#include "kernels.h"
#include "path.h"

namespace cr {

// ------------------------------------------------------------------ helpers
__device__ inline uint32_t wave_append(uint32_t* counter, bool pred) {
    const unsigned long long mask = __ballot(pred);
    if (!mask) return 0xffffffffu;
    const int lane = __lane_id();
    const int leader = __ffsll((long long)mask) - 1;
    uint32_t base = 0;
    if (lane == leader) base = atomicAdd(counter, uint32_t(__popcll(mask)));
    base = __shfl(base, leader);
    return base + uint32_t(__popcll(mask & ((1ull << lane) - 1ull)));
}

__device__ inline unsigned long long wave_claim(unsigned long long* counter, bool pred) {
    const unsigned long long mask = __ballot(pred);
    if (!mask) return ~0ull;
    const int lane = __lane_id();
    const int leader = __ffsll((long long)mask) - 1;
    unsigned long long base = 0;
    if (lane == leader) base = atomicAdd(counter, (unsigned long long)__popcll(mask));
    base = __shfl(base, leader);
    return base + (unsigned long long)__popcll(mask & ((1ull << lane) - 1ull));
}

// Sum a per-lane counter over the wave and add it once to `dst`.
__device__ inline void wave_add(unsigned long long* dst, uint32_t v) {
    unsigned long long s = v;
    for (int o = 32; o > 0; o >>= 1) s += __shfl_xor(s, o);
    if (__lane_id() == 0 && s) atomicAdd(dst, s);
}

__device__ inline void report_error(Control* ctl, uint32_t job, uint64_t key, uint32_t code, uint32_t depth,
                                    uint32_t tri, uint32_t material, sf::f64 t, V3 point, V3 origin, V3 dir) {
    const uint32_t slot = atomicAdd(&ctl->err_count, 1u);
    if (slot >= uint32_t(kMaxErrors)) return;
    ErrorRec& e = ctl->errors[slot];
    e.key = key;
    e.job = job;
    e.code = code;
    e.depth = depth;
    e.triangle = tri;
    e.material = material;
    e.tbits = t.b;
    e.point[0] = point.x; e.point[1] = point.y; e.point[2] = point.z;
    e.origin[0] = origin.x; e.origin[1] = origin.y; e.origin[2] = origin.z;
    e.direction[0] = dir.x; e.direction[1] = dir.y; e.direction[2] = dir.z;
}

__device__ inline V3 xyz(float4 v) { return V3{v.x, v.y, v.z}; }

// ------------------------------------------------------------------ ray queries
// Bvh::intersect over an array of rays (parity harness).  One lane per ray.
__global__ void __launch_bounds__(256) k_trace_rays(const DevScene* scene, const RayIn* rays, uint64_t n,
                                                    HitOut* hits, int any, Control* ctl) {
    const uint64_t i = uint64_t(blockIdx.x) * blockDim.x + threadIdx.x;
    if (i >= n) return;
    const DevScene& s = *scene;
    const RayIn r = rays[i];
    RayQuery q;
    setup_query(q, V3{r.o[0], r.o[1], r.o[2]}, V3{r.d[0], r.d[1], r.d[2]}, r.tmin, r.tmax,
                V3{r.lo[0], r.lo[1], r.lo[2]}, r.ignore);
    TraceCount tc{0, 0, 0};
    ClosestResult best = any ? traverse<true>(s, q, tc) : traverse<false>(s, q, tc);
    HitOut h;
    h.hit = best.prim != kNoTriangle ? 1u : 0u;
    h.triangle = h.hit ? best.id : kNoTriangle;
    h.t = h.u = h.v = 0;
    if (best.overflow) h.hit = 2u;
    if (h.hit == 1u) {
        if (!best.exact) {
            const ExactHit e = exact_of(s, q, best.prim);
            best.t = e.t;
            best.u = e.u;
            best.v = e.v;
        }
        h.t = best.t.b;
        h.u = best.u.b;
        h.v = best.v.b;
    }
    hits[i] = h;
    if (tc.fallbacks) atomicAdd(&ctl->fallbacks, (unsigned long long)tc.fallbacks);
}

// ------------------------------------------------------------------ path state
__device__ inline uint32_t find_job(const DevJob* jobs, uint32_t njobs, uint64_t item) {
    uint32_t lo = 0, hi = njobs;
    while (hi - lo > 1) {
        const uint32_t mid = (lo + hi) >> 1;
        if (jobs[mid].item_base <= item) lo = mid;
        else hi = mid;
    }
    return lo;
}

__device__ inline PathState load_state(const PathSoA& P, uint32_t slot) {
    PathState st;
    const float4 ro = P.ray_o[slot], rd = P.ray_d[slot], rl = P.ray_lo[slot];
    const float4 th = P.thr[slot], ra = P.rad[slot], lp = P.lastpos[slot], ab = P.absorb[slot];
    st.o = xyz(ro);
    st.tmin = ro.w;
    st.d = xyz(rd);
    st.tmax = rd.w;
    st.lo = xyz(rl);
    st.T = xyz(th);
    st.eta_scale = th.w;
    st.L = xyz(ra);
    st.last_pdf = ra.w;
    st.last_pos = xyz(lp);
    st.sigma = xyz(ab);
    st.medium = P.medium[slot];
    st.flags = P.flags[slot];
    st.prev = P.prev[slot];
    st.rng = P.rng[slot];
    return st;
}

__device__ inline void store_state(const PathSoA& P, uint32_t slot, const PathState& st) {
    P.ray_o[slot] = make_float4(st.o.x, st.o.y, st.o.z, st.tmin);
    P.ray_d[slot] = make_float4(st.d.x, st.d.y, st.d.z, st.tmax);
    P.ray_lo[slot] = make_float4(st.lo.x, st.lo.y, st.lo.z, 0);
    P.thr[slot] = make_float4(st.T.x, st.T.y, st.T.z, st.eta_scale);
    P.rad[slot] = make_float4(st.L.x, st.L.y, st.L.z, st.last_pdf);
    P.lastpos[slot] = make_float4(st.last_pos.x, st.last_pos.y, st.last_pos.z, 0);
    P.absorb[slot] = make_float4(st.sigma.x, st.sigma.y, st.sigma.z, 0);
    P.medium[slot] = st.medium;
    P.flags[slot] = st.flags;
    P.prev[slot] = st.prev;
    P.rng[slot] = st.rng;
}

// Camera sample `local` of job `ji` into `slot`; false if the camera failed.
__device__ inline bool init_path(const PathSoA& P, uint32_t slot, const DevJob& J, uint32_t ji, uint64_t local,
                                 bool precise, Control* ctl) {
    const uint64_t npix = uint64_t(J.width) * uint64_t(J.height);
    const uint32_t sample = uint32_t(local / npix);
    const uint32_t pixel = uint32_t(local - uint64_t(sample) * npix);
    PathState st;
    const bool ok = begin_sample(J, pixel, sample, precise, st);
    P.job[slot] = ji;
    P.pixel[slot] = pixel;
    P.sample[slot] = sample;
    store_state(P, slot, st);
    if (!ok)
        report_error(ctl, ji, uint64_t(pixel) * uint64_t(J.spp) + sample, kErrCamera, 0, 0, 0, sf::zero(0),
                     V3{0, 0, 0}, V3{0, 0, 0}, V3{0, 0, 0});
    return ok;
}

__global__ void __launch_bounds__(256) k_seed(PathSoA P, const DevJob* jobs, uint32_t njobs, uint32_t pool,
                                              uint32_t* q0, Control* ctl) {
    const uint32_t slot = blockIdx.x * blockDim.x + threadIdx.x;
    bool queued = false;
    if (slot < pool && uint64_t(slot) < ctl->total) {
        const uint32_t j = find_job(jobs, njobs, slot);
        queued = init_path(P, slot, jobs[j], j, uint64_t(slot) - jobs[j].item_base, false, ctl);
    }
    const uint32_t pos = wave_append(&ctl->qcount[0], queued);
    if (queued) q0[pos] = slot;
}

// Exact-queue indexing: consecutive entries go to different waves, so a wave
// holds as few undecided rays as the grid allows.  Emulated binary64 runs per
// lane with divergent control flow; packing those rays into few waves would
// serialise them.
__device__ inline uint32_t spread_waves() { return gridDim.x * (blockDim.x / 64); }
__device__ inline uint32_t spread_first() {
    return (blockIdx.x * (blockDim.x / 64) + threadIdx.x / 64) + spread_waves() * (threadIdx.x % 64);
}
__device__ inline uint32_t spread_stride() { return spread_waves() * 64; }

// ------------------------------------------------------------------ extend
// FP32-only closest hit; rays with an undecided triangle go to the exact queue.
__global__ void __launch_bounds__(256) k_extend(PathSoA P, const DevJob* jobs, const uint32_t* queue, uint32_t* xq,
                                                int parity, Control* ctl) {
    const uint32_t n = ctl->qcount[parity];
    const uint32_t stride = gridDim.x * blockDim.x;
    const uint32_t rounds = (n + stride - 1) / stride;
    TraceCount tc{0, 0, 0};
    uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
    for (uint32_t r = 0; r < rounds; ++r, i += stride) {
        bool exact = false;
        uint32_t path = 0;
        if (i < n) {
            path = queue[i];
            const DevScene& s = *jobs[P.job[path]].scene;
            const float4 ro = P.ray_o[path], rd = P.ray_d[path], rl = P.ray_lo[path];
            RayQuery q;
            setup_query(q, xyz(ro), xyz(rd), ro.w, rd.w, xyz(rl), P.prev[path]);
            const ClosestResult best = traverse<false, false>(s, q, tc);
            exact = best.need_exact && !best.overflow;
            P.hit[path] = best.overflow ? kNoTriangle - 1 : best.prim;
        }
        const uint32_t pos = wave_append(&ctl->xcount, exact);
        if (exact) xq[pos] = path;
    }
    wave_add(&ctl->extend_nodes, tc.nodes);
    wave_add(&ctl->extend_tris, tc.tris);
    if (blockIdx.x == 0 && threadIdx.x == 0) atomicAdd(&ctl->extend_rays, (unsigned long long)n);
}

// Closest hit with binary64-exact resolution of undecided tests.
__global__ void __launch_bounds__(256) k_extend_exact(PathSoA P, const DevJob* jobs, const uint32_t* xq, Control* ctl) {
    const uint32_t n = ctl->xcount;
    TraceCount tc{0, 0, 0};
    for (uint32_t i = spread_first(); i < n; i += spread_stride()) {
        const uint32_t path = xq[i];
        const DevScene& s = *jobs[P.job[path]].scene;
        const float4 ro = P.ray_o[path], rd = P.ray_d[path], rl = P.ray_lo[path];
        RayQuery q;
        setup_query(q, xyz(ro), xyz(rd), ro.w, rd.w, xyz(rl), P.prev[path]);
        const ClosestResult best = traverse<false, true>(s, q, tc);
        P.hit[path] = best.overflow ? kNoTriangle - 1 : best.prim;
    }
    wave_add(&ctl->fallbacks, tc.fallbacks);
    if (blockIdx.x == 0 && threadIdx.x == 0) atomicAdd(&ctl->exact_rays, (unsigned long long)n);
}

// ------------------------------------------------------------------ shade
__global__ void __launch_bounds__(256) k_shade(PathSoA P, ShadowSoA S, const DevJob* jobs, const uint32_t* queue,
                                               uint32_t* next_queue, uint32_t* retire, int parity, Control* ctl) {
    const uint32_t n = ctl->qcount[parity];
    const uint32_t stride = gridDim.x * blockDim.x;
    const uint32_t rounds = (n + stride - 1) / stride;
    uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
    for (uint32_t r = 0; r < rounds; ++r, i += stride) {
        const bool live = i < n;
        uint32_t path = 0, action = kRetire;
        ShadowOut sh;
        sh.emit = false;
        if (live) {
            path = queue[i];
            const uint32_t ji = P.job[path];
            const DevJob& J = jobs[ji];
            const uint64_t key = uint64_t(P.pixel[path]) * uint64_t(J.spp) + P.sample[path];
            const uint32_t prim = P.hit[path];
            PathState st = load_state(P, path);
            if (prim == kNoTriangle - 1) {
                report_error(ctl, ji, key, kErrStack, st.flags & kDepthMask, 0, 0, sf::zero(0), st.o, st.o, st.d);
                st.L = v3(0);
                action = kRetire;
            } else {
                ErrorOut err;
                action = bounce(J, st, prim, sh, err);
                if (err.code)
                    report_error(ctl, ji, key, err.code, err.depth, err.triangle, err.material, err.t, err.point,
                                 err.origin, err.direction);
            }
            if (action == kReplay) {
                atomicAdd(&ctl->replays, 1ull);
                begin_sample(J, P.pixel[path], P.sample[path], true, st);
                action = kContinue;
            }
            store_state(P, path, st);
        }
        const uint32_t spos = wave_append(&ctl->scount, live && sh.emit);
        if (live && sh.emit) {
            S.path[spos] = path;
            S.o[spos] = make_float4(sh.o.x, sh.o.y, sh.o.z, 0.0f);
            S.d[spos] = make_float4(sh.d.x, sh.d.y, sh.d.z, sh.tmax);
            S.lo[spos] = make_float4(sh.lo.x, sh.lo.y, sh.lo.z, __uint_as_float(sh.ignore));
            S.c[spos] = make_float4(sh.c.x, sh.c.y, sh.c.z, 0);
        }
        const uint32_t qpos = wave_append(&ctl->qcount[parity ^ 1], live && action == kContinue);
        if (live && action == kContinue) next_queue[qpos] = path;
        const uint32_t rpos = wave_append(&ctl->rcount, live && action == kRetire);
        if (live && action == kRetire) retire[rpos] = path;
    }
}

// ------------------------------------------------------------------ shadow
__device__ inline void add_contribution(const PathSoA& P, const ShadowSoA& S, uint32_t i) {
    const uint32_t path = S.path[i];
    const float4 c = S.c[i];
    float4 r = P.rad[path];
    r.x = r.x + c.x;
    r.y = r.y + c.y;
    r.z = r.z + c.z;
    P.rad[path] = r;
}

__device__ inline void shadow_query(RayQuery& q, const ShadowSoA& S, uint32_t i) {
    const float4 so = S.o[i], sd = S.d[i], sl = S.lo[i];
    setup_query(q, xyz(so), xyz(sd), so.w, sd.w, xyz(sl), __float_as_uint(sl.w));
}

// FP32-only visibility; undecided segments without a certain blocker go to the exact queue.
__global__ void __launch_bounds__(256) k_shadow(PathSoA P, ShadowSoA S, const DevJob* jobs, uint32_t* xsq,
                                                Control* ctl) {
    const uint32_t n = ctl->scount;
    const uint32_t stride = gridDim.x * blockDim.x;
    const uint32_t rounds = (n + stride - 1) / stride;
    TraceCount tc{0, 0, 0};
    uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
    for (uint32_t r = 0; r < rounds; ++r, i += stride) {
        bool exact = false;
        if (i < n) {
            const DevScene& s = *jobs[P.job[S.path[i]]].scene;
            RayQuery q;
            shadow_query(q, S, i);
            const ClosestResult b = traverse<true, false>(s, q, tc);
            if (b.prim == kNoTriangle && !b.overflow) {
                if (b.need_exact) exact = true;
                else add_contribution(P, S, i);
            }
        }
        const uint32_t pos = wave_append(&ctl->xscount, exact);
        if (exact) xsq[pos] = i;
    }
    wave_add(&ctl->shadow_nodes, tc.nodes);
    wave_add(&ctl->shadow_tris, tc.tris);
    if (blockIdx.x == 0 && threadIdx.x == 0) atomicAdd(&ctl->shadow_rays, (unsigned long long)n);
}

__global__ void __launch_bounds__(256) k_shadow_exact(PathSoA P, ShadowSoA S, const DevJob* jobs, const uint32_t* xsq,
                                                      Control* ctl) {
    const uint32_t n = ctl->xscount;
    TraceCount tc{0, 0, 0};
    for (uint32_t j = spread_first(); j < n; j += spread_stride()) {
        const uint32_t i = xsq[j];
        const DevScene& s = *jobs[P.job[S.path[i]]].scene;
        RayQuery q;
        shadow_query(q, S, i);
        const ClosestResult b = traverse<true, true>(s, q, tc);
        if (b.prim == kNoTriangle && !b.overflow) add_contribution(P, S, i);
    }
    wave_add(&ctl->fallbacks, tc.fallbacks);
    if (blockIdx.x == 0 && threadIdx.x == 0) atomicAdd(&ctl->exact_rays, (unsigned long long)n);
}

// ------------------------------------------------------------------ retire + regenerate
__global__ void __launch_bounds__(256) k_retire(PathSoA P, const DevJob* jobs, uint32_t njobs, const uint32_t* retire,
                                                uint32_t* next_queue, int parity, Control* ctl) {
    const uint32_t n = ctl->rcount;
    const uint32_t stride = gridDim.x * blockDim.x;
    const uint32_t rounds = (n + stride - 1) / stride;
    uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
    for (uint32_t r = 0; r < rounds; ++r, i += stride) {
        const bool live = i < n;
        uint32_t path = 0;
        if (live) {
            path = retire[i];
            const uint32_t ji = P.job[path];
            const DevJob& J = jobs[ji];
            const float4 L = P.rad[path];
            const V3 l = xyz(L);
            const uint64_t npix = uint64_t(J.width) * uint64_t(J.height);
            if (!radiance_ok(l)) {
                report_error(ctl, ji, uint64_t(P.pixel[path]) * uint64_t(J.spp) + P.sample[path], kErrBadRadiance,
                             0, 0, 0, sf::zero(0), V3{0, 0, 0}, V3{0, 0, 0}, V3{0, 0, 0});
            }
            J.samples[uint64_t(P.sample[path]) * npix + P.pixel[path]] = make_float4(l.x, l.y, l.z, 0);
        }
        const unsigned long long item = wave_claim(&ctl->issued, live);
        bool queued = false;
        if (live && item < ctl->total) {
            const uint32_t j = find_job(jobs, njobs, item);
            queued = init_path(P, path, jobs[j], j, item - jobs[j].item_base, false, ctl);
        }
        if (live) atomicAdd(&ctl->retired, 1ull);
        const uint32_t qpos = wave_append(&ctl->qcount[parity ^ 1], queued);
        if (queued) next_queue[qpos] = path;
    }
}

__global__ void k_reset(int parity, Control* ctl) {
    ctl->qcount[parity] = 0;
    ctl->scount = 0;
    ctl->rcount = 0;
    ctl->xcount = 0;
    ctl->xscount = 0;
}

// ------------------------------------------------------------------ accumulate
// film = float(sum_s L_s / spp) with the sum in binary64, samples in order.
__global__ void __launch_bounds__(256) k_accumulate(const DevJob* jobs, uint32_t job, float* film) {
    const DevJob& J = jobs[job];
    const uint64_t npix = uint64_t(J.width) * uint64_t(J.height);
    const uint64_t p = uint64_t(blockIdx.x) * blockDim.x + threadIdx.x;
    if (p >= npix) return;
    sf::f64 r = sf::zero(0), g = sf::zero(0), b = sf::zero(0);
    for (int s = 0; s < J.spp; ++s) {
        const float4 v = J.samples[uint64_t(s) * npix + p];
        r = sf::add(r, sf::from_float(v.x));
        g = sf::add(g, sf::from_float(v.y));
        b = sf::add(b, sf::from_float(v.z));
    }
    const sf::f64 n = sf::from_i32(J.spp);
    film[3 * p + 0] = sf::to_float(sf::div(r, n));
    film[3 * p + 1] = sf::to_float(sf::div(g, n));
    film[3 * p + 2] = sf::to_float(sf::div(b, n));
}

// ------------------------------------------------------------------ launchers
static inline unsigned blocks_for(uint64_t n, unsigned bs) { return unsigned((n + bs - 1) / bs); }

hipError_t launch_trace_rays(const DevScene* scene, const RayIn* rays, uint64_t n, HitOut* hits, int any, Control* ctl,
                             hipStream_t st) {
    if (n == 0) return hipSuccess;
    hipLaunchKernelGGL(k_trace_rays, dim3(blocks_for(n, 256)), dim3(256), 0, st, scene, rays, n, hits, any, ctl);
    return hipGetLastError();
}

hipError_t launch_seed(const PathSoA& P, const DevJob* jobs, uint32_t njobs, uint32_t pool, uint32_t* q0, Control* ctl,
                       hipStream_t st) {
    hipLaunchKernelGGL(k_seed, dim3(blocks_for(pool, 256)), dim3(256), 0, st, P, jobs, njobs, pool, q0, ctl);
    return hipGetLastError();
}

hipError_t launch_stage(Stage stage, const PathSoA& P, const ShadowSoA& S, const DevJob* jobs, uint32_t njobs,
                        const Queues& Q, int parity, Control* ctl, unsigned grid, hipStream_t st) {
    switch (stage) {
        case kStageExtend:
            hipLaunchKernelGGL(k_extend, dim3(grid), dim3(256), 0, st, P, jobs, Q.q[parity], Q.xq, parity, ctl);
            break;
        case kStageExtendExact:
            hipLaunchKernelGGL(k_extend_exact, dim3(grid), dim3(256), 0, st, P, jobs, Q.xq, ctl);
            break;
        case kStageShade:
            hipLaunchKernelGGL(k_shade, dim3(grid), dim3(256), 0, st, P, S, jobs, Q.q[parity], Q.q[parity ^ 1],
                               Q.retire, parity, ctl);
            break;
        case kStageShadow:
            hipLaunchKernelGGL(k_shadow, dim3(grid), dim3(256), 0, st, P, S, jobs, Q.xsq, ctl);
            break;
        case kStageShadowExact:
            hipLaunchKernelGGL(k_shadow_exact, dim3(grid), dim3(256), 0, st, P, S, jobs, Q.xsq, ctl);
            break;
        case kStageRetire:
            hipLaunchKernelGGL(k_retire, dim3(grid), dim3(256), 0, st, P, jobs, njobs, Q.retire, Q.q[parity ^ 1],
                               parity, ctl);
            break;
        default:
            hipLaunchKernelGGL(k_reset, dim3(1), dim3(1), 0, st, parity, ctl);
            break;
    }
    return hipGetLastError();
}

hipError_t launch_accumulate(const DevJob* jobs, uint32_t job, uint64_t npix, float* film, hipStream_t st) {
    hipLaunchKernelGGL(k_accumulate, dim3(blocks_for(npix, 256)), dim3(256), 0, st, jobs, job, film);
    return hipGetLastError();
}

}  // namespace cr
