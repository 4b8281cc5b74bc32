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

#include <algorithm>
#include <map>
#include <mutex>

#include "kernels.h"
#include "film.h"
#include "path.h"

// Minimum workgroups of 256 lanes per compute unit the compiler must fit
// (register budget = occupancy).  Traversal: 4 (128 VGPRs, 4 waves per SIMD)
// measured fastest on the 3.1 M and 8.0 M triangle renders, ahead of 5 (96
// VGPRs, heavy spilling) and 3 (no spills, too few waves to hide node-fetch
// latency).  Shading: 3 (168 VGPRs, some spilling in the rare binary64
// chains) measured 0.9 % faster than 2 (256 VGPRs) on the 3.1 M triangle
// render and the same on the 8.0 M one; 4 spills on the common path (-13 %).
#ifndef CR_TRAVERSE_MIN_BLOCKS
#define CR_TRAVERSE_MIN_BLOCKS 4
#endif
#ifndef CR_SHADE_MIN_BLOCKS
#define CR_SHADE_MIN_BLOCKS 3
#endif

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

// Sum a per-lane counter over the wave and add it once to `dst`.
__device__ inline void wave_add(unsigned long long* dst, uint32_t v) {
    unsigned long long s = v;
    for (int o = 32; o > 0; o >>= 1) s += __shfl_xor(s, o);
    if (__lane_id() == 0 && s) atomicAdd(dst, s);
}

// Sum `v` per distinct `key` over the wave; one atomic per key into arr[key].
// The loop is wave-uniform and runs once per distinct key (usually once).
__device__ inline void wave_add_keyed(unsigned long long* arr, uint32_t key, uint32_t v) {
    unsigned long long pending = __ballot(v != 0);
    const int lane = __lane_id();
    while (pending) {
        const int leader = __ffsll((long long)pending) - 1;
        const uint32_t k = __shfl(key, leader);
        const bool mine = ((pending >> lane) & 1ull) && key == k;
        pending &= ~__ballot(mine);
        unsigned long long s = mine ? v : 0;
        for (int o = 32; o > 0; o >>= 1) s += __shfl_xor(s, o);
        if (lane == leader) atomicAdd(&arr[k], s);
    }
}

// Per-job counters of one workgroup, summed in LDS over the whole launch and
// added to the device counters once at its end: one atomic per job and
// workgroup instead of one per wave and round (same-address device atomics
// from every XCD serialise).
struct WorkgroupCounts {
    unsigned long long* slots;  // kJobSlots entries in LDS
    __device__ void clear() const {
        for (uint32_t k = threadIdx.x; k < kJobSlots; k += blockDim.x) slots[k] = 0;
    }
    __device__ void flush(unsigned long long* dst) const {
        for (uint32_t k = threadIdx.x; k < kJobSlots; k += blockDim.x)
            if (slots[k]) atomicAdd(&dst[k], slots[k]);
    }
};

// Errors go to a ring the host drains after every step; a step that raises
// more than kMaxErrors errors keeps the last kMaxErrors records.
__device__ inline void report_error(Control* ctl, uint32_t job, uint64_t key, uint32_t code, uint32_t depth,
                                    uint32_t tri, uint32_t material, sf::f64 t, V3 point, V3 origin, V3 dir) {
    const uint32_t n = atomicAdd(&ctl->err_count, 1u);
    ErrorRec& e = ctl->errors[n % uint32_t(kMaxErrors)];
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
template <int W>
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
    ClosestResult best = any ? traverse<true, true, W>(s, q, tc) : traverse<false, true, W>(s, q, tc);
    HitOut h;
    h.hit = best.prim != kNoTriangle ? 1u : 0u;
    h.triangle = h.hit ? best.id : kNoTriangle;
    h.t = h.u = h.v = 0;
    if (best.overflow) h.hit = 2u;
    if (best.tex_error) h.hit = 3u;
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
// Slot of the job owning global work item `item`: the last live sequence
// number whose item_base is <= item (item_base grows with the sequence).
__device__ inline uint32_t find_job(const DevJob* jobs, const Control* ctl, uint64_t item) {
    uint32_t lo = ctl->seq_lo, hi = ctl->seq_hi;
    while (hi - lo > 1) {
        const uint32_t mid = lo + ((hi - lo) >> 1);
        if (jobs[mid % kJobSlots].item_base <= item) lo = mid;
        else hi = mid;
    }
    return lo % kJobSlots;
}

__device__ inline PathState load_state(const PathSoA& P, uint32_t slot) {
    PathState st;
    const PathRay& r = P.ray[slot];
    const PathRec& c = P.rec[slot];
    const float4 ro = r.o, rd = r.d, rl = r.lo;
    const float4 th = c.thr, ra = c.rad, lp = c.lastpos, ab = c.absorb;
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
    st.medium = c.medium;
    st.flags = r.flags;
    st.prev = r.prev;
    st.rng = c.rng;
    return st;
}

// Writes the ray record (except job and hit) and the state record (except
// pixel and sample).
__device__ inline void store_state(const PathSoA& P, uint32_t slot, const PathState& st) {
    PathRay& r = P.ray[slot];
    PathRec& c = P.rec[slot];
    r.o = make_float4(st.o.x, st.o.y, st.o.z, st.tmin);
    r.d = make_float4(st.d.x, st.d.y, st.d.z, st.tmax);
    r.lo = make_float4(st.lo.x, st.lo.y, st.lo.z, 0);
    r.prev = st.prev;
    r.flags = st.flags;
    c.thr = make_float4(st.T.x, st.T.y, st.T.z, st.eta_scale);
    c.rad = make_float4(st.L.x, st.L.y, st.L.z, st.last_pdf);
    c.lastpos = make_float4(st.last_pos.x, st.last_pos.y, st.last_pos.z, 0);
    c.absorb = make_float4(st.sigma.x, st.sigma.y, st.sigma.z, 0);
    c.medium = st.medium;
    c.rng = st.rng;
}

// Camera sample `local` of job `ji` into `slot`; false if the camera failed.
__device__ inline bool init_path(const PathSoA& P, uint32_t slot, const DevJob& J, uint32_t ji, uint64_t local,
                                 bool precise, Control* ctl) {
    const uint64_t npix = uint64_t(J.win_w) * uint64_t(J.win_h);
    const uint32_t sample = uint32_t(local / npix);
    const uint32_t lp = uint32_t(local - uint64_t(sample) * npix);  // window pixel
    const uint32_t ly = lp / uint32_t(J.win_w), lx = lp - ly * uint32_t(J.win_w);
    const uint32_t pixel = (uint32_t(J.win_y) + ly) * uint32_t(J.width) + uint32_t(J.win_x) + lx;
    PathState st;
    const bool ok = begin_sample(J, pixel, sample, precise, st);
    if (!ok) {  // retired by the next shade with zero radiance
        st.flags |= kFailed;
        st.L = V3{0, 0, 0};
        st.tmin = 0;
        st.tmax = -1;
    }
    P.ray[slot].job = ji;
    P.rec[slot].pixel = pixel;
    P.rec[slot].sample = sample;
    store_state(P, slot, st);
    if (!ok)
        report_error(ctl, ji, uint64_t(pixel) * uint64_t(J.spp) + sample, kErrCamera, 0, 0, 0, sf::zero(0),
                     V3{0, 0, 0}, V3{0, 0, 0}, V3{0, 0, 0});
    return ok;
}

// Start work item `item` in path slot `path` (a failed camera ray is still
// queued: shade retires it with zero radiance after the error is recorded).
// Inlined into the retire and refill kernels (k_retire measured 3 % faster
// than with a call).
__attribute__((always_inline)) __device__ inline void start_item(const PathSoA& P, const DevJob* jobs, uint32_t path, uint64_t item,
                                  Control* ctl) {
    const uint32_t job = find_job(jobs, ctl, item);
    const DevJob& J = jobs[job];
    if (item >= J.item_base && item - J.item_base < J.items) {
        init_path(P, path, J, job, item - J.item_base, false, ctl);
        return;
    }
    // Unreachable with a consistent job table; kept so that a bookkeeping bug
    // surfaces as an error instead of a write outside a sample buffer.
    report_error(ctl, job, 0, kErrInternal, 0, 0, 0, sf::zero(0), V3{0, 0, 0}, V3{0, 0, 0}, V3{0, 0, 0});
    PathRay& r = P.ray[path];
    r.job = job;
    r.flags = kFailed;
    r.o = make_float4(0, 0, 0, 0);
    r.d = make_float4(0, 0, 1, -1);
    r.lo = make_float4(0, 0, 0, 0);
    r.prev = kNoTriangle;
    P.rec[path].pixel = 0;
    P.rec[path].sample = uint32_t(J.spp);  // outside the sample buffer: never stored
    P.rec[path].rad = make_float4(0, 0, 0, 0);
}

// ------------------------------------------------------------------ extend
// Persistent FP32-only closest hit.  Each lane runs one query at a time as a
// resumable traversal; every few visits, when a quarter of the wave is idle,
// the idle lanes claim the next queue entries together (one atomic per wave),
// so a wave's time follows the mean query length rather than its longest
// query.  Rays with an undecided triangle go to the exact queue.
#ifndef CR_VISITS
#define CR_VISITS 16
#endif
constexpr int kVisitsPerRefill = CR_VISITS;
#ifndef CR_REFILL_BELOW
#define CR_REFILL_BELOW 48
#endif
constexpr int kRefillBelow = CR_REFILL_BELOW;  // refill when at most this many lanes are busy

// kVisitsPerRefill rounds of traversal for the busy lanes of a wave.  With
// CR_TRAV_PP each round is one inner-node visit per lane, followed by a leaf
// round for the lanes holding a leaf when no lane is still looking for its
// first leaf or at least CR_LEAF_MIN lanes hold one; otherwise each round is
// one node or leaf visit per lane.
#ifndef CR_TRAV_PP
#define CR_TRAV_PP 1
#endif
#ifndef CR_LEAF_MIN
#define CR_LEAF_MIN 32
#endif
template <class T>
__device__ inline void run_visits(T& tr, const DevScene* sc, RayQuery& q, TraceCount& tc, bool& busy, bool& fin) {
#if CR_TRAV_PP
    for (int k = 0; k < kVisitsPerRefill; ++k) {
        if (busy && tr.inner(*sc, q, tc)) {
            busy = false;
            fin = true;
        }
        if (!__ballot(busy)) break;
        const bool searching = busy && tr.leaf == kEmptyChild && tr.can_descend();
        const bool ready = busy && tr.has_leaf();
        if (!__ballot(searching) || __popcll(__ballot(ready)) >= CR_LEAF_MIN) {
            if (ready && tr.leaves(*sc, q, tc)) {
                busy = false;
                fin = true;
            }
        }
    }
#else
    for (int k = 0; k < kVisitsPerRefill && busy; ++k)
        if (tr.step(*sc, q, tc)) {
            busy = false;
            fin = true;
        }
#endif
}

// Newest traversal stack entries of each lane in LDS (lane-minor layout, so a
// wave's accesses to one ring slot are conflict-free); older entries overflow
// to the lane's private array.  18 entries x 8 B x 256 lanes = 36 KiB per
// workgroup, 144 KiB of the 160 KiB LDS at 4 workgroups per CU.
#ifndef CR_LDS_RING
#define CR_LDS_RING 18
#endif
// The ring pointers carry the LDS address space explicitly: through generic
// pointers the accesses would compile to flat instructions.
using lds_u32 = __attribute__((address_space(3))) uint32_t;
using lds_f32 = __attribute__((address_space(3))) float;
// Rows are kRingLanes entries apart whatever the workgroup size (at most
// 256 threads, one column each): a compile-time row stride makes a slot's
// address a shift instead of a 32-bit multiply (quarter rate) per access.
constexpr uint32_t kRingLanes = 256;
template <int N>
struct LdsRing {
    static constexpr int kSize = N;
    lds_u32* ref;  // this lane's column: ref[i * kRingLanes]
    lds_f32* near;
    __device__ void put(int i, uint32_t r, float n) {
        ref[i * kRingLanes] = r;
        near[i * kRingLanes] = n;
    }
    __device__ void get(int i, uint32_t& r, float& n) const {
        r = ref[i * kRingLanes];
        n = near[i * kRingLanes];
    }
};
using TraverseRing = LdsRing<CR_LDS_RING>;

__device__ inline uint32_t wave_claim32(uint32_t* counter, bool pred) {
    const unsigned long long mask = __ballot(pred);
    if (!mask) return 0xffffffffu;
    const int lane = __lane_id();
    const int leader = __ffsll((long long)mask) - 1;
    uint32_t base = 0;
    if (lane == leader) base = atomicAdd(counter, uint32_t(__popcll(mask)));
    base = __shfl(base, leader);
    return base + uint32_t(__popcll(mask & ((1ull << lane) - 1ull)));
}

// Claim of queue entries for the idle lanes of a wave from a queue of n
// entries split into kClaimParts contiguous parts, each with its own head
// (heads[p * 32]): one atomic per claim on the wave's current part; a lane
// whose entry would fall past the part's end gets none, and the wave moves to
// the next part once its part is used up.  `tried` counts the parts the wave
// has found used up: all of them (tried == kClaimParts) means the queue is
// drained for this wave.  Returns the lane's queue index or 0xffffffff.
__device__ inline uint32_t wave_claim_parted(uint32_t* heads, uint32_t n, uint32_t& part, uint32_t& tried,
                                             bool pred) {
    const unsigned long long mask = __ballot(pred);
    if (!mask || tried >= kClaimParts) return 0xffffffffu;
    const int lane = __lane_id();
    const int leader = __ffsll((long long)mask) - 1;
    const uint32_t want = uint32_t(__popcll(mask));
    const uint32_t lo = uint32_t(uint64_t(part) * n / kClaimParts);
    const uint32_t hi = uint32_t(uint64_t(part + 1) * n / kClaimParts);
    uint32_t base = 0;
    if (lane == leader) base = atomicAdd(&heads[part * 32], want);
    base = __shfl(base, leader);
    if (base + want >= hi - lo || base >= hi - lo) {  // this part is used up (for every wave)
        part = part + 1 == kClaimParts ? 0 : part + 1;
        ++tried;
    }
    const uint32_t local = base + uint32_t(__popcll(mask & ((1ull << lane) - 1ull)));
    return pred && local < hi - lo ? lo + local : 0xffffffffu;
}

// Closest-hit word stored per path: primitive index in bits 0..29 (or one of
// two codes), shading class in bits 30..31 so that shade can group paths by
// class without touching the scene.
constexpr uint32_t kHitPrimMask = 0x3fffffffu;
constexpr uint32_t kHitMiss = 0x3fffffffu;      // no hit: environment
constexpr uint32_t kHitOverflow = 0x3ffffffeu;  // traversal stack exhausted
constexpr uint32_t kHitTexture = 0x3ffffffdu;   // alpha-mask texture coordinates not finite

__device__ inline uint32_t encode_hit(const DevScene& s, const ClosestResult& b) {
    if (b.overflow) return kHitOverflow;
    if (b.tex_error) return kHitTexture;
    if (b.prim == kNoTriangle) return kHitMiss;
    return b.prim | (((s.prims[b.prim].flags >> kPrimClassShift) & 3u) << 30);
}

// Both traversal kernels come in two variants, with and without the FP32
// alpha-mask tests (MASK); the one that does not match ctl->masks returns at
// once, so an iteration runs the mask variant only while a live job needs it.
// COUNT adds the per-query work counters (wide nodes visited, triangles
// tested); the timed variant carries no counter state in its loop.
template <bool MASK, bool COUNT, int W>
__global__ void __launch_bounds__(256, CR_TRAVERSE_MIN_BLOCKS) k_extend(PathSoA P, const DevJob* jobs, const uint32_t* queue, uint32_t* xq,
                                                Pending* xpend, int parity, Control* ctl, JobCounters* cnt) {
    if ((ctl->masks != 0) != MASK) return;
    const uint32_t n = ctl->qcount[parity];
    __shared__ uint32_t ring_ref[CR_LDS_RING * kRingLanes];
    __shared__ float ring_near[CR_LDS_RING * kRingLanes];
    StackEntry overflow[kStackSize];
    Pending cands;
    Traversal<false, false, TraverseRing, (CR_LEAF_PRELOAD != 0), MASK, W> tr;
    tr.stack = overflow;
    tr.cands = &cands;
    tr.ring = TraverseRing{(lds_u32*)(ring_ref + threadIdx.x), (lds_f32*)(ring_near + threadIdx.x)};
    RayQuery q;
    const DevScene* sc = nullptr;
    uint32_t path = 0, job = 0;
    bool busy = false, drained = false;
    uint32_t part = blockIdx.x % kClaimParts, tried = 0;
    TraceCount tc{0, 0, 0};
    for (;;) {
        if (!drained && __popcll(__ballot(busy)) <= kRefillBelow) {
            const uint32_t idx = wave_claim_parted(ctl->eparts, n, part, tried, !busy);
            if (tried >= kClaimParts) drained = true;
            if (!busy && idx < n) {
                path = queue[idx];
                const PathRay r = P.ray[path];
                job = r.job;
                sc = &jobs[job].scene;
                setup_query(q, xyz(r.o), xyz(r.d), r.o.w, r.d.w, xyz(r.lo), r.prev);
                tc = TraceCount{0, 0, 0};
                busy = !tr.begin(*sc, q);
                if (!busy) P.ray[path].hit = kHitMiss;
            }
        }
        if (!__ballot(busy)) {
            if (drained) break;
            continue;
        }
        bool fin = false;
        run_visits(tr, sc, q, tc, busy, fin);
        const int work = fin ? exact_work<false>(tr.best) : 0;
        if (fin && !work) P.ray[path].hit = encode_hit(*sc, tr.best);
        const uint32_t pos = wave_append(&ctl->xcount, work != 0);
        if (work) {
            xq[pos] = path;
            xpend[pos] = pending_of(tr.best, work);
        }
        if constexpr (COUNT) {
            wave_add_keyed(cnt->extend_nodes, job, fin ? tc.nodes : 0u);
            wave_add_keyed(cnt->extend_tris, job, fin ? tc.tris : 0u);
        }
    }
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

// Closest hit with binary64-exact resolution of what the FP32 pass left.
__global__ void __launch_bounds__(256) k_extend_exact(PathSoA P, const DevJob* jobs, const uint32_t* xq,
                                                      const Pending* xpend, Control* ctl, JobCounters* cnt) {
    const uint32_t n = ctl->xcount;
    for (uint32_t i = spread_first(); i < n; i += spread_stride()) {
        const uint32_t path = xq[i];
        const PathRay r = P.ray[path];
        const uint32_t job = r.job;
        const DevScene& s = jobs[job].scene;
        RayQuery q;
        setup_query(q, xyz(r.o), xyz(r.d), r.o.w, r.d.w, xyz(r.lo), r.prev);
        TraceCount tc{0, 0, 0};
        const ClosestResult best = exact_pass<false>(s, q, xpend[i], tc);
        P.ray[path].hit = encode_hit(s, best);
        atomicAdd(&cnt->fallbacks[job], (unsigned long long)tc.fallbacks);
        atomicAdd(&cnt->exact_rays[job], 1ull);
    }
}

// ------------------------------------------------------------------ shade
// Shade in chunks of kShadeChunk queue entries per workgroup: the chunk is
// first counting-sorted by shading class in LDS, so each wave shades one kind
// of surface (or only misses) instead of paying for every branch.
// 2048 entries (eight per thread) measured 0.8 % faster than 1024 on the
// 3.1 M triangle render at three shading workgroups per CU, the same on the
// 8.0 M one.
#ifndef CR_SHADE_CHUNK
#define CR_SHADE_CHUNK 2048
#endif
constexpr uint32_t kShadeChunk = CR_SHADE_CHUNK;
#ifndef CR_SHADE_OCTSORT
#define CR_SHADE_OCTSORT 1
#endif

// One path's bounce given its closest-hit word: the state update, the shadow
// ray it emits and what becomes of the path.
struct ShadeResult {
    uint32_t action, oct;
    bool replayed;
    ShadowOut sh;
};

__device__ inline ShadeResult shade_one(const PathSoA& P, const DevJob& J, uint32_t ji, uint32_t path,
                                        uint32_t hit_word, Control* ctl) {
    ShadeResult o;
    o.replayed = false;
    o.sh.emit = false;
    const uint32_t pixel = P.rec[path].pixel, sample = P.rec[path].sample;
    const uint64_t key = uint64_t(pixel) * uint64_t(J.spp) + sample;
    const uint32_t hit = hit_word & kHitPrimMask;
    PathState st = load_state(P, path);
    if (st.flags & kFailed) {
        o.action = kRetire;
    } else if (hit == kHitOverflow || hit == kHitTexture) {
        report_error(ctl, ji, key, hit == kHitOverflow ? kErrStack : kErrTexture, st.flags & kDepthMask, 0, 0,
                     sf::zero(0), st.o, st.o, st.d);
        st.L = v3(0);
        o.action = kRetire;
    } else {
        ErrorOut err;
        o.action = bounce(J, st, hit == kHitMiss ? kNoTriangle : hit, o.sh, err);
        if (err.code)
            report_error(ctl, ji, key, err.code, err.depth, err.triangle, err.material, err.t, err.point, err.origin,
                         err.direction);
    }
    if (o.action == kReplay) {
        o.replayed = true;
        begin_sample(J, pixel, sample, true, st);
        o.action = kContinue;
    }
    if (o.action == kRetire) {
        // A finished path's records are read again only for its radiance
        // (k_shadow adds an unoccluded last contribution, k_retire takes the
        // sum); the slot's next start rewrites both records whole.
        P.rec[path].rad = make_float4(st.L.x, st.L.y, st.L.z, st.last_pdf);
    } else {
        store_state(P, path, st);
    }
    o.oct = (st.d.x < 0 ? 1u : 0u) | (st.d.y < 0 ? 2u : 0u) | (st.d.z < 0 ? 4u : 0u);
    return o;
}

// Wave-level bookkeeping of shade_one's results: counters, the shadow queue
// and the retire list (the continuing paths' next-queue slot is the caller's).
__device__ inline void shade_emit(const ShadowSoA& S, const WorkgroupCounts* wc, Control* ctl, uint32_t* retire,
                                  uint32_t* retire_job, bool live, uint32_t path, uint32_t ji, const ShadeResult& o) {
    wave_add_keyed(wc[0].slots, ji, live ? 1u : 0u);
    wave_add_keyed(wc[1].slots, ji, live && o.sh.emit ? 1u : 0u);
    wave_add_keyed(wc[2].slots, ji, live && o.replayed ? 1u : 0u);
    const uint32_t spos = wave_append(&ctl->scount, live && o.sh.emit);
    if (live && o.sh.emit) {
        S.path[spos] = path;
        S.job[spos] = ji;
        S.o[spos] = make_float4(o.sh.o.x, o.sh.o.y, o.sh.o.z, 0.0f);
        S.d[spos] = make_float4(o.sh.d.x, o.sh.d.y, o.sh.d.z, o.sh.tmax);
        S.lo[spos] = make_float4(o.sh.lo.x, o.sh.lo.y, o.sh.lo.z, __uint_as_float(o.sh.ignore));
        S.c[spos] = make_float4(o.sh.c.x, o.sh.c.y, o.sh.c.z, 0);
    }
    const uint32_t rpos = wave_append(&ctl->rcount, live && o.action == kRetire);
    if (live && o.action == kRetire) {
        retire[rpos] = path;
        retire_job[rpos] = ji;
    }
}

__global__ void __launch_bounds__(256, CR_SHADE_MIN_BLOCKS) k_shade(PathSoA P, ShadowSoA S, const DevJob* jobs,
                                                                  const uint32_t* queue, uint32_t* next_queue,
                                                                  uint32_t* retire, uint32_t* retire_job, int parity,
                                                                  Control* ctl, JobCounters* cnt) {
    __shared__ uint32_t stage[kShadeChunk], sorted[kShadeChunk];
    __shared__ uint32_t counts[kShadeClasses], cursor[kShadeClasses];
    __shared__ unsigned long long wg_extend[kJobSlots], wg_shadow[kJobSlots], wg_replays[kJobSlots];
    const WorkgroupCounts wc[3] = {{wg_extend}, {wg_shadow}, {wg_replays}};
    for (const WorkgroupCounts& c : wc) c.clear();
#if CR_SHADE_OCTSORT
    __shared__ uint32_t ocount[8], ocursor[8], obase;
#endif
    const uint32_t n = ctl->qcount[parity];
    // Entries per thread and chunk: kShadeChunk / blockDim.x while the queue
    // fills the grid that many times over; a shorter queue (the last
    // iterations of a render, when only the longest paths are left) is cut
    // into smaller chunks so that it spreads over more workgroups instead of
    // being shaded by a few threads in turn (same results: the chunk only
    // orders the work).
    const uint32_t lanes = gridDim.x * blockDim.x;
    const uint32_t per = min(kShadeChunk / blockDim.x, max(1u, (n + lanes - 1) / lanes));
    const uint32_t chunk = per * blockDim.x;
    // Workgroups claim chunks one at a time from a queue counter: only two
    // shading workgroups fit a CU, and with a fixed grid-stride assignment the
    // last round of chunks left most CUs idle (k_shade 4.8 % slower).
    __shared__ uint32_t chunk_base;
    for (;;) {
        if (threadIdx.x < kShadeClasses) counts[threadIdx.x] = cursor[threadIdx.x] = 0;
        if (threadIdx.x == 0) chunk_base = atomicAdd(&ctl->khead, chunk);
        __syncthreads();
        const uint32_t base = chunk_base;  // the same for the whole workgroup
        if (base >= n) break;
        const uint32_t m = min(chunk, n - base);
        for (uint32_t k = threadIdx.x; k < m; k += blockDim.x) {  // path ids < 2^25: class fits above
            const uint32_t path = queue[base + k];
            const uint32_t c = P.ray[path].hit >> 30;
            stage[k] = path | (c << 30);
            atomicAdd(&counts[c], 1u);
        }
        __syncthreads();
        for (uint32_t k = threadIdx.x; k < m; k += blockDim.x) {
            const uint32_t v = stage[k], c = v >> 30;
            uint32_t off = 0;
            for (uint32_t j = 0; j < c; ++j) off += counts[j];
            sorted[off + atomicAdd(&cursor[c], 1u)] = v & kHitPrimMask;
        }
#if CR_SHADE_OCTSORT
        if (threadIdx.x < 8) ocount[threadIdx.x] = 0;
#endif
        __syncthreads();
        for (uint32_t r = 0; r < per; ++r) {
            const uint32_t k = r * blockDim.x + threadIdx.x;
            const bool live = k < m;
            uint32_t path = 0, ji = 0;
            ShadeResult o;
            o.action = kRetire;
            o.oct = 0;
            o.replayed = false;
            o.sh.emit = false;
            if (live) {
                path = sorted[k];
                const uint32_t hit_word = P.ray[path].hit;
                ji = P.ray[path].job;
                o = shade_one(P, jobs[ji], ji, path, hit_word, ctl);
            }
            shade_emit(S, wc, ctl, retire, retire_job, live, path, ji, o);
#if CR_SHADE_OCTSORT
            if (live) {
                const bool cont = o.action == kContinue;
                stage[k] = cont ? (path | (o.oct << 29)) : 0xffffffffu;
                if (cont) atomicAdd(&ocount[o.oct], 1u);
            }
#else
            const uint32_t qpos = wave_append(&ctl->qcount[parity ^ 1], live && o.action == kContinue);
            if (live && o.action == kContinue) next_queue[qpos] = path;
#endif
        }
#if CR_SHADE_OCTSORT
        // The chunk's continuing paths go to the next extend queue grouped by
        // direction octant, so the traversal's waves hold rays of one octant.
        __syncthreads();
        if (threadIdx.x == 0) {
            uint32_t t = 0;
            for (int o = 0; o < 8; ++o) {
                ocursor[o] = t;
                t += ocount[o];
            }
            obase = t ? atomicAdd(&ctl->qcount[parity ^ 1], t) : 0u;
        }
        __syncthreads();
        for (uint32_t k = threadIdx.x; k < m; k += blockDim.x) {
            const uint32_t v = stage[k];
            if (v == 0xffffffffu) continue;
            next_queue[obase + atomicAdd(&ocursor[v >> 29], 1u)] = v & 0x1fffffffu;
        }
#endif
        __syncthreads();  // `sorted` is rewritten by the next chunk
    }
    wc[0].flush(cnt->extend_rays);
    wc[1].flush(cnt->shadow_rays);
    wc[2].flush(cnt->replays);
}

// ------------------------------------------------------------------ shadow
__device__ inline void add_contribution(const PathSoA& P, const ShadowSoA& S, uint32_t i) {
    const uint32_t path = S.path[i];
    const float4 c = S.c[i];
    float4 r = P.rec[path].rad;
    r.x = r.x + c.x;
    r.y = r.y + c.y;
    r.z = r.z + c.z;
    P.rec[path].rad = r;
}

__device__ inline void shadow_query(RayQuery& q, const ShadowSoA& S, uint32_t i) {
    const float4 so = S.o[i], sd = S.d[i], sl = S.lo[i];
    setup_query(q, xyz(so), xyz(sd), so.w, sd.w, xyz(sl), __float_as_uint(sl.w));
}

// Persistent FP32-only visibility (same scheme as k_extend); undecided
// segments without a certain blocker go to the exact queue.
template <bool MASK, bool COUNT, int W>
__global__ void __launch_bounds__(256, CR_TRAVERSE_MIN_BLOCKS) k_shadow(PathSoA P, ShadowSoA S, const DevJob* jobs, uint32_t* xsq,
                                                Pending* xspend, Control* ctl, JobCounters* cnt) {
    if ((ctl->masks != 0) != MASK) return;
    const uint32_t n = ctl->scount;
    __shared__ uint32_t ring_ref[CR_LDS_RING * kRingLanes];
    __shared__ float ring_near[CR_LDS_RING * kRingLanes];
    StackEntry overflow[kStackSize];
    Pending cands;
    Traversal<true, false, TraverseRing, (CR_LEAF_PRELOAD != 0), MASK, W> tr;
    tr.stack = overflow;
    tr.cands = &cands;
    tr.ring = TraverseRing{(lds_u32*)(ring_ref + threadIdx.x), (lds_f32*)(ring_near + threadIdx.x)};
    RayQuery q;
    const DevScene* sc = nullptr;
    uint32_t i = 0, job = 0;
    bool busy = false, drained = false;
    uint32_t part = blockIdx.x % kClaimParts, tried = 0;
    TraceCount tc{0, 0, 0};
    for (;;) {
        bool fin = false;
        if (!drained && __popcll(__ballot(busy)) <= kRefillBelow) {
            const uint32_t idx = wave_claim_parted(ctl->sparts, n, part, tried, !busy);
            if (tried >= kClaimParts) drained = true;
            if (!busy && idx < n) {
                i = idx;
                job = S.job[i];
                sc = &jobs[job].scene;
                shadow_query(q, S, i);
                tc = TraceCount{0, 0, 0};
                busy = !tr.begin(*sc, q);
                fin = !busy;  // nothing to traverse: unoccluded
            }
        }
        if (!__ballot(busy || fin)) {
            if (drained) break;
            continue;
        }
        run_visits(tr, sc, q, tc, busy, fin);
        int work = 0;
        if (fin && tr.best.prim == kNoTriangle && !tr.best.overflow) {
            work = exact_work<true>(tr.best);
            if (!work) add_contribution(P, S, i);
        }
        const uint32_t pos = wave_append(&ctl->xscount, work != 0);
        if (work) {
            xsq[pos] = i;
            xspend[pos] = pending_of(tr.best, work);
        }
        if constexpr (COUNT) {
            wave_add_keyed(cnt->shadow_nodes, job, fin ? tc.nodes : 0u);
            wave_add_keyed(cnt->shadow_tris, job, fin ? tc.tris : 0u);
        }
    }
}

__global__ void __launch_bounds__(256) k_shadow_exact(PathSoA P, ShadowSoA S, const DevJob* jobs, const uint32_t* xsq,
                                                      const Pending* xspend, Control* ctl, JobCounters* cnt) {
    const uint32_t n = ctl->xscount;
    for (uint32_t j = spread_first(); j < n; j += spread_stride()) {
        const uint32_t i = xsq[j];
        const uint32_t job = S.job[i];
        const DevScene& s = jobs[job].scene;
        RayQuery q;
        shadow_query(q, S, i);
        TraceCount tc{0, 0, 0};
        const ClosestResult b = exact_pass<true>(s, q, xspend[j], tc);
        if (b.tex_error) {
            const uint32_t path = S.path[i];
            const uint64_t key = uint64_t(P.rec[path].pixel) * uint64_t(jobs[job].spp) + P.rec[path].sample;
            report_error(ctl, job, key, kErrTexture, 0, 0, 0, sf::zero(0), q.o, q.o, q.d);
        } else if (b.prim == kNoTriangle && !b.overflow) {
            add_contribution(P, S, i);
        }
        atomicAdd(&cnt->fallbacks[job], (unsigned long long)tc.fallbacks);
        atomicAdd(&cnt->exact_rays[job], 1ull);
    }
}

// ------------------------------------------------------------------ retire + regenerate
// Finished paths store their sample.  Retire-queue entry i takes work item
// issued + i while items are left (and joins the next extend queue after the
// paths shade continued), otherwise its slot goes onto the free stack.  All
// positions follow from i and the counters as k_retire found them, so no
// atomics are needed; k_reset then advances the counters.
__device__ inline uint32_t retire_takes(const Control* ctl) {
    const unsigned long long issued = ctl->issued, total = ctl->total;
    const unsigned long long avail = issued < total ? total - issued : 0ull;
    return avail < ctl->rcount ? uint32_t(avail) : ctl->rcount;
}

__global__ void __launch_bounds__(256) k_retire(PathSoA P, const DevJob* jobs, const uint32_t* retire,
                                                const uint32_t* retire_job,
                                                uint32_t* next_queue, uint32_t* freeq, int parity, Control* ctl,
                                                JobCounters* cnt) {
    const uint32_t n = ctl->rcount;
    const uint32_t take = retire_takes(ctl);
    const unsigned long long issued = ctl->issued;
    const uint32_t q0 = ctl->qcount[parity ^ 1], f0 = ctl->fcount;
    __shared__ unsigned long long wg_retired[kJobSlots];
    const WorkgroupCounts wc{wg_retired};
    wc.clear();
    __syncthreads();
    const uint32_t stride = gridDim.x * blockDim.x;
    const uint32_t rounds = (n + stride - 1) / stride;
    uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
    for (uint32_t r = 0; r < rounds; ++r, i += stride) {
        const bool live = i < n;
        uint32_t path = 0, ji = 0;
        if (live) {
            path = retire[i];
            ji = retire_job[i];
            const DevJob& J = jobs[ji];
            const PathRec& c = P.rec[path];
            const V3 l = xyz(c.rad);
            const uint32_t pixel = c.pixel, sample = c.sample;
            const uint64_t npix = uint64_t(J.win_w) * uint64_t(J.win_h);
            const uint32_t gy = pixel / uint32_t(J.width), gx = pixel - gy * uint32_t(J.width);
            const uint64_t lp = uint64_t(gy - uint32_t(J.win_y)) * uint64_t(J.win_w) + (gx - uint32_t(J.win_x));
            if (!radiance_ok(l)) {
                report_error(ctl, ji, uint64_t(pixel) * uint64_t(J.spp) + sample, kErrBadRadiance, 0, 0, 0,
                             sf::zero(0), V3{0, 0, 0}, V3{0, 0, 0}, V3{0, 0, 0});
            }
            const uint64_t at = uint64_t(sample) * npix + lp;
            if (lp < npix && at < J.items) J.samples[at] = make_float4(l.x, l.y, l.z, 0);
        }
        wave_add_keyed(wc.slots, ji, live ? 1u : 0u);
        if (live) {
            if (i < take) {
                start_item(P, jobs, path, issued + i, ctl);
                next_queue[q0 + i] = path;
            } else {
                freeq[f0 + (i - take)] = path;
            }
        }
    }
    __syncthreads();
    wc.flush(cnt->retired);
}

// End of an iteration: apply k_retire's claims, clear the consumed counters
// and reserve the work items that idle slots start in the following refill.
// Idempotent at a step boundary (rcount and the refill range are then zero
// or already accounted), so the host may run it again before a refill.
__global__ void k_reset(int parity, Control* ctl) {
    const uint32_t take = retire_takes(ctl);
    ctl->qcount[parity ^ 1] += take;
    ctl->fcount += ctl->rcount - take;
    ctl->issued += take;
    ctl->qcount[parity] = 0;
    ctl->scount = 0;
    ctl->rcount = 0;
    ctl->xcount = 0;
    ctl->xscount = 0;
    ctl->ehead = 0;
    ctl->shead = 0;
    for (uint32_t p = 0; p < kClaimParts; ++p) ctl->eparts[p * 32] = ctl->sparts[p * 32] = 0;
    ctl->khead = 0;
    const unsigned long long issued = ctl->issued, total = ctl->total;
    const unsigned long long avail = issued < total ? total - issued : 0ull;
    const uint32_t refill = avail < ctl->fcount ? uint32_t(avail) : ctl->fcount;
    ctl->refill_base = issued;
    ctl->refill_n = refill;
    ctl->fcount -= refill;
    ctl->refill_from = ctl->fcount;
    ctl->refill_q = ctl->qcount[parity ^ 1];
    ctl->qcount[parity ^ 1] += refill;
    ctl->issued = issued + refill;
}

// Idle slots start the items reserved by k_reset; they join the next extend queue.
__global__ void __launch_bounds__(256) k_refill(PathSoA P, const DevJob* jobs, const uint32_t* freeq,
                                                uint32_t* next_queue, int parity, Control* ctl) {
    const uint32_t n = ctl->refill_n;
    const uint32_t stride = gridDim.x * blockDim.x;
    const uint32_t rounds = (n + stride - 1) / stride;
    uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
    for (uint32_t r = 0; r < rounds; ++r, i += stride) {
        const bool live = i < n;
        uint32_t path = 0;
        if (live) {
            path = freeq[ctl->refill_from + i];
            start_item(P, jobs, path, ctl->refill_base + i, ctl);
        }
        if (live) next_queue[ctl->refill_q + i] = path;
    }
}

// Host-side admission: new total and live sequence range; the admitted job's
// slot starts with zero counters.
__global__ void k_admit(Control* ctl, JobCounters* cnt, unsigned long long total, uint32_t seq_lo, uint32_t seq_hi,
                        uint32_t slot) {
    ctl->total = total;
    ctl->seq_lo = seq_lo;
    ctl->seq_hi = seq_hi;
    cnt->retired[slot] = 0;
    cnt->replays[slot] = 0;
    cnt->extend_rays[slot] = 0;
    cnt->shadow_rays[slot] = 0;
    cnt->extend_nodes[slot] = 0;
    cnt->extend_tris[slot] = 0;
    cnt->shadow_nodes[slot] = 0;
    cnt->shadow_tris[slot] = 0;
    cnt->fallbacks[slot] = 0;
    cnt->exact_rays[slot] = 0;
}

// Initial control state for a pool of `pool` idle slots (all on the free stack).
__global__ void __launch_bounds__(256) k_pool_init(uint32_t* freeq, uint32_t pool, Control* ctl) {
    const uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i < pool) freeq[i] = i;
    if (i == 0) {
        ctl->qcount[0] = ctl->qcount[1] = 0;
        ctl->scount = ctl->rcount = ctl->xcount = ctl->xscount = 0;
        ctl->fcount = pool;
        ctl->refill_n = ctl->refill_from = ctl->refill_q = 0;
        ctl->seq_lo = ctl->seq_hi = 0;
        ctl->err_count = 0;
        ctl->ehead = ctl->shead = ctl->khead = 0;
        for (uint32_t p = 0; p < kClaimParts; ++p) ctl->eparts[p * 32] = ctl->sparts[p * 32] = 0;
        ctl->pool = pool;
        ctl->issued = ctl->total = ctl->refill_base = 0;
        ctl->fallbacks = 0;
    }
}

// ------------------------------------------------------------------ accumulate
// film = float(sum_s L_s / spp) with the sum in binary64, samples in order.
__global__ void __launch_bounds__(256) k_accumulate(const DevJob* jobs, uint32_t job, float* film) {
    const DevJob& J = jobs[job];
    const uint64_t npix = uint64_t(J.win_w) * uint64_t(J.win_h);
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
                             int width, hipStream_t st) {
    if (n == 0) return hipSuccess;
    if (width == 8)
        hipLaunchKernelGGL(k_trace_rays<8>, dim3(blocks_for(n, 256)), dim3(256), 0, st, scene, rays, n, hits, any, ctl);
    else
        hipLaunchKernelGGL(k_trace_rays<4>, dim3(blocks_for(n, 256)), dim3(256), 0, st, scene, rays, n, hits, any, ctl);
    return hipGetLastError();
}

// ------------------------------------------------------------------ per-function probe
__device__ inline Surf surf_of(const crg_surface& c) {
    Surf s;
    s.position = V3{c.position[0], c.position[1], c.position[2]};
    s.position_low = V3{c.position_low[0], c.position_low[1], c.position_low[2]};
    s.ng = V3{c.geometric_normal[0], c.geometric_normal[1], c.geometric_normal[2]};
    s.normal = V3{c.normal[0], c.normal[1], c.normal[2]};
    s.base = V3{c.base_color[0], c.base_color[1], c.base_color[2]};
    s.emission = V3{c.emission[0], c.emission[1], c.emission[2]};
    s.att_color = V3{c.attenuation_color[0], c.attenuation_color[1], c.attenuation_color[2]};
    s.metallic = c.metallic;
    s.roughness = c.roughness;
    s.transmission = c.transmission;
    s.ior = c.ior;
    s.thickness = c.thickness;
    s.att_distance = c.attenuation_distance;
    s.material = c.material;
    s.front = c.front_face != 0;
    s.double_sided = c.double_sided != 0;
    s.texture_error = false;
    return s;
}

__device__ inline void put3(float* o, V3 v) {
    o[0] = v.x;
    o[1] = v.y;
    o[2] = v.z;
}

__device__ inline crg_probe_ray probe_ray(V3 o, V3 d, float tmin, float tmax, int32_t status) {
    crg_probe_ray r{};
    put3(r.ray.origin, o);
    put3(r.ray.direction, d);
    r.ray.tmin = tmin;
    r.ray.tmax = tmax;
    r.ray.ignore = 0xffffffffu;
    r.status = status;
    return r;
}

__global__ void __launch_bounds__(256) k_probe(ProbeArgs a) {
    const uint64_t i = uint64_t(blockIdx.x) * blockDim.x + threadIdx.x;
    if (i >= a.n) return;
    switch (a.what) {
        case CRG_PROBE_CAMERA: {
            const float* px = static_cast<const float*>(a.in) + 2 * i;
            V3 o, d;
            float tmin = 0, tmax = 0;
            const bool ok = camera_ray(a.camera, px[0], px[1], o, d, tmin, tmax);
            static_cast<crg_probe_ray*>(a.out)[i] = probe_ray(o, d, tmin, tmax, ok ? 0 : 1);
            break;
        }
        case CRG_PROBE_ENV_SAMPLE: {
            Rng rng{static_cast<const uint64_t*>(a.in)[i]};
            crg_probe_env_sample r{};
            put3(r.direction, env_sample(*a.env, rng, r.pdf));
            r.rng_end = rng.state;
            static_cast<crg_probe_env_sample*>(a.out)[i] = r;
            break;
        }
        case CRG_PROBE_ENV_PDF: {
            const float* d = static_cast<const float*>(a.in) + 3 * i;
            static_cast<float*>(a.out)[i] = env_pdf(*a.env, V3{d[0], d[1], d[2]});
            break;
        }
        case CRG_PROBE_ENV_RADIANCE: {
            const float* d = static_cast<const float*>(a.in) + 3 * i;
            put3(static_cast<float*>(a.out) + 3 * i, env_radiance(*a.env, V3{d[0], d[1], d[2]}));
            break;
        }
        case CRG_PROBE_BSDF_SAMPLE: {
            const crg_probe_bsdf_in& q = static_cast<const crg_probe_bsdf_in*>(a.in)[i];
            Rng rng{q.rng};
            const BsdfSample b = sample_bsdf(surf_of(q.surface), V3{q.wo[0], q.wo[1], q.wo[2]}, rng);
            crg_probe_bsdf_sample r{};
            put3(r.direction, b.direction);
            put3(r.weight, b.weight);
            r.pdf = b.pdf;
            r.eta = b.eta;
            r.delta = b.delta;
            r.transmitted = b.transmitted;
            r.valid = sample_ok(b);
            r.rng_end = rng.state;
            static_cast<crg_probe_bsdf_sample*>(a.out)[i] = r;
            break;
        }
        case CRG_PROBE_BSDF_EVAL: {
            const crg_probe_bsdf_in& q = static_cast<const crg_probe_bsdf_in*>(a.in)[i];
            const BsdfEval e = evaluate_bsdf(surf_of(q.surface), V3{q.wo[0], q.wo[1], q.wo[2]},
                                             V3{q.wi[0], q.wi[1], q.wi[2]});
            float* o = static_cast<float*>(a.out) + 4 * i;
            put3(o, e.value);
            o[3] = e.pdf;
            break;
        }
        case CRG_PROBE_OFFSET_RAY: {
            const crg_probe_offset_in& q = static_cast<const crg_probe_offset_in*>(a.in)[i];
            V3 o{0, 0, 0}, d{0, 0, 0};
            float tmax = 0;
            const uint32_t err = offset_ray(V3{q.position[0], q.position[1], q.position[2]},
                                            V3{q.geometric_normal[0], q.geometric_normal[1], q.geometric_normal[2]},
                                            V3{a.ung[3 * i], a.ung[3 * i + 1], a.ung[3 * i + 2]},
                                            V3{q.direction[0], q.direction[1], q.direction[2]}, q.max_distance, o, d, tmax);
            static_cast<crg_probe_ray*>(a.out)[i] = probe_ray(o, d, 0.0f, tmax, int32_t(err));
            break;
        }
        default: {  // CRG_PROBE_SURFACE_AT
            const crg_probe_surface_in& q = static_cast<const crg_probe_surface_in*>(a.in)[i];
            const uint32_t prim = a.tri_to_prim[q.hit.triangle];
            // The position as shading computes it (float-float where that
            // decides its rounding), the low part from the binary64 chain.
            const V3 dir{q.ray.direction[0], q.ray.direction[1], q.ray.direction[2]};
            const sf::f64 hu{__builtin_bit_cast(uint64_t, q.hit.u)}, hv{__builtin_bit_cast(uint64_t, q.hit.v)};
            const Surf s = surface_at(*a.scene, prim, dir, hu, hv, false);
            const Surf sl = surface_at(*a.scene, prim, dir, hu, hv, true);
            crg_surface r{};
            put3(r.position, s.position);
            put3(r.position_low, sl.position_low);
            put3(r.geometric_normal, s.ng);
            put3(r.normal, s.normal);
            put3(r.base_color, s.base);
            put3(r.emission, s.emission);
            put3(r.attenuation_color, s.att_color);
            r.metallic = s.metallic;
            r.roughness = s.roughness;
            r.transmission = s.transmission;
            r.ior = s.ior;
            r.thickness = s.thickness;
            r.attenuation_distance = s.att_distance;
            r.material = s.material;
            r.front_face = s.front;
            r.double_sided = s.double_sided;
            static_cast<crg_surface*>(a.out)[i] = r;
            break;
        }
    }
}

hipError_t launch_probe(const ProbeArgs& a, hipStream_t st) {
    if (a.n == 0) return hipSuccess;
    hipLaunchKernelGGL(k_probe, dim3(blocks_for(a.n, 256)), dim3(256), 0, st, a);
    return hipGetLastError();
}

// ------------------------------------------------------------------ allocation-free ray queries
// Fixed device scratch of crg_trace_rays_noalloc: the control block, then the
// staged rays and hits of at most kNoAllocMaxRays queries.  Every device the
// library runs on holds its own copy for the life of the process; the host
// reaches it only through the address hipGetSymbolAddress returns for that
// device (noalloc_scratch), never through the symbol's name.
static __device__ __align__(256) uint8_t scratch[kNoAllocScratchBytes];

__global__ void k_clear_words(uint32_t* p, uint32_t n) {
    for (uint32_t i = blockIdx.x * blockDim.x + threadIdx.x; i < n; i += gridDim.x * blockDim.x) p[i] = 0;
}

hipError_t launch_clear_words(uint32_t* p, uint32_t n, hipStream_t st) {
    hipLaunchKernelGGL(k_clear_words, dim3(std::min<uint32_t>(64, (n + 255) / 256)), dim3(256), 0, st, p, n);
    return hipGetLastError();
}

hipError_t noalloc_scratch(int device, void** out) {
    static std::mutex mu;
    static std::map<int, void*> addr;
    std::lock_guard<std::mutex> lock(mu);
    auto it = addr.find(device);
    if (it == addr.end()) {
        int current = -1;
        hipError_t e = hipGetDevice(&current);
        if (e != hipSuccess) return e;
        if (current != device) return hipErrorInvalidDevice;  // the caller makes `device` current first
        void* p = nullptr;
        e = hipGetSymbolAddress(&p, HIP_SYMBOL(scratch));
        if (e != hipSuccess) return e;
        it = addr.emplace(device, p).first;
    }
    *out = it->second;
    return hipSuccess;
}

// Both alpha-mask variants of a traversal stage (the one not matching
// ctl->masks returns at once), with or without work counters, for W-wide nodes.
template <int W, bool SHADOW>
void launch_traversal(bool count, dim3 g, dim3 b, hipStream_t st, const PathSoA& P, const ShadowSoA& S,
                      const DevJob* jobs, const Queues& Q, int parity, Control* ctl, JobCounters* cnt) {
    if constexpr (SHADOW) {
        if (count) {
            hipLaunchKernelGGL((k_shadow<false, true, W>), g, b, 0, st, P, S, jobs, Q.xsq, Q.xspend, ctl, cnt);
            hipLaunchKernelGGL((k_shadow<true, true, W>), g, b, 0, st, P, S, jobs, Q.xsq, Q.xspend, ctl, cnt);
        } else {
            hipLaunchKernelGGL((k_shadow<false, false, W>), g, b, 0, st, P, S, jobs, Q.xsq, Q.xspend, ctl, cnt);
            hipLaunchKernelGGL((k_shadow<true, false, W>), g, b, 0, st, P, S, jobs, Q.xsq, Q.xspend, ctl, cnt);
        }
    } else {
        if (count) {
            hipLaunchKernelGGL((k_extend<false, true, W>), g, b, 0, st, P, jobs, Q.q[parity], Q.xq, Q.xpend, parity,
                               ctl, cnt);
            hipLaunchKernelGGL((k_extend<true, true, W>), g, b, 0, st, P, jobs, Q.q[parity], Q.xq, Q.xpend, parity,
                               ctl, cnt);
        } else {
            hipLaunchKernelGGL((k_extend<false, false, W>), g, b, 0, st, P, jobs, Q.q[parity], Q.xq, Q.xpend, parity,
                               ctl, cnt);
            hipLaunchKernelGGL((k_extend<true, false, W>), g, b, 0, st, P, jobs, Q.q[parity], Q.xq, Q.xpend, parity,
                               ctl, cnt);
        }
    }
}

hipError_t launch_stage(Stage stage, const PathSoA& P, const ShadowSoA& S, const DevJob* jobs, const Queues& Q,
                        int parity, Control* ctl, JobCounters* cnt, LaunchShape shape, hipStream_t st) {
    const dim3 g(shape.grid), b(shape.block);
    switch (stage) {
        case kStageExtend:
            if (shape.width == 8) launch_traversal<8, false>(shape.count, g, b, st, P, S, jobs, Q, parity, ctl, cnt);
            else launch_traversal<4, false>(shape.count, g, b, st, P, S, jobs, Q, parity, ctl, cnt);
            break;
        case kStageExtendExact:
            hipLaunchKernelGGL(k_extend_exact, g, b, 0, st, P, jobs, Q.xq, Q.xpend, ctl, cnt);
            break;
        case kStageShade:
            hipLaunchKernelGGL(k_shade, g, b, 0, st, P, S, jobs, Q.q[parity], Q.q[parity ^ 1], Q.retire, Q.retire_job,
                               parity, ctl, cnt);
            break;
        case kStageShadow:
            if (shape.width == 8) launch_traversal<8, true>(shape.count, g, b, st, P, S, jobs, Q, parity, ctl, cnt);
            else launch_traversal<4, true>(shape.count, g, b, st, P, S, jobs, Q, parity, ctl, cnt);
            break;
        case kStageShadowExact:
            hipLaunchKernelGGL(k_shadow_exact, g, b, 0, st, P, S, jobs, Q.xsq, Q.xspend, ctl, cnt);
            break;
        case kStageRetire:
            hipLaunchKernelGGL(k_retire, g, b, 0, st, P, jobs, Q.retire, Q.retire_job, Q.q[parity ^ 1], Q.freeq, parity,
                               ctl, cnt);
            break;
        case kStageReset:
            hipLaunchKernelGGL(k_reset, dim3(1), dim3(1), 0, st, parity, ctl);
            break;
        default:
            hipLaunchKernelGGL(k_refill, g, b, 0, st, P, jobs, Q.freeq, Q.q[parity ^ 1], parity, ctl);
            break;
    }
    return hipGetLastError();
}

hipError_t launch_pool_init(uint32_t* freeq, uint32_t pool, Control* ctl, hipStream_t st) {
    hipLaunchKernelGGL(k_pool_init, dim3(blocks_for(pool, 256)), dim3(256), 0, st, freeq, pool, ctl);
    return hipGetLastError();
}

__global__ void k_set_masks(Control* ctl, uint32_t masks) { ctl->masks = masks; }

hipError_t launch_set_masks(Control* ctl, uint32_t masks, hipStream_t st) {
    hipLaunchKernelGGL(k_set_masks, dim3(1), dim3(1), 0, st, ctl, masks);
    return hipGetLastError();
}

hipError_t launch_admit(Control* ctl, JobCounters* cnt, uint64_t total, uint32_t seq_lo, uint32_t seq_hi, uint32_t slot,
                        hipStream_t st) {
    hipLaunchKernelGGL(k_admit, dim3(1), dim3(1), 0, st, ctl, cnt, (unsigned long long)total, seq_lo, seq_hi, slot);
    return hipGetLastError();
}

// ------------------------------------------------------------------ film encode
// One lane per pixel: validity (lowest failing pixel per condition), three
// display bytes by threshold search in LDS, and the RGBE bytes into the row's
// four planes.
__global__ void __launch_bounds__(256) k_film_encode(const float* film, uint64_t npix, uint32_t width,
                                                     const float* thresholds, uint8_t* ldr, uint8_t* rgbe,
                                                     uint32_t* first_bad) {
    __shared__ float t[kDisplayLevels - 1];
    if (ldr)
        for (uint32_t k = threadIdx.x; k < kDisplayLevels - 1; k += blockDim.x) t[k] = thresholds[k];
    __syncthreads();
    const uint64_t i = uint64_t(blockIdx.x) * blockDim.x + threadIdx.x;
    if (i >= npix) return;
    const float r = film[3 * i], g = film[3 * i + 1], b = film[3 * i + 2];
    if (!finite1(r) || !finite1(g) || !finite1(b) || r < 0 || g < 0 || b < 0) atomicMin(&first_bad[0], uint32_t(i));
    else if (fmaxf(r, fmaxf(g, b)) >= 1.7014118346046923e38f) atomicMin(&first_bad[1], uint32_t(i));  // 2^127
    if (ldr) {
        ldr[3 * i] = uint8_t(display_byte(r, t));
        ldr[3 * i + 1] = uint8_t(display_byte(g, t));
        ldr[3 * i + 2] = uint8_t(display_byte(b, t));
    }
    if (rgbe) {
        uint8_t q[4];
        rgbe_of(r, g, b, q);
        const uint64_t row = i / width, x = i - row * width;
        uint8_t* base = rgbe + row * 4ull * width + x;
#pragma unroll
        for (int k = 0; k < 4; ++k) base[uint64_t(k) * width] = q[k];
    }
}

// One lane per (row, plane): the plane's run-length code, written to its own
// fixed-size slot.
__global__ void __launch_bounds__(256) k_film_rle(const uint8_t* rgbe, uint32_t width, uint32_t height,
                                                  uint8_t* streams, uint32_t* lens) {
    const uint64_t j = uint64_t(blockIdx.x) * blockDim.x + threadIdx.x;
    if (j >= uint64_t(height) * 4u) return;
    const uint8_t* in = rgbe + j * width;  // plane j % 4 of row j / 4
    lens[j] = rle_component(in, 1u, width, streams + j * rle_bound(width));
}

hipError_t launch_film_encode(const float* film, uint64_t npix, uint32_t width, const float* thresholds, uint8_t* ldr,
                              uint8_t* rgbe, uint32_t* first_bad, hipStream_t st) {
    hipLaunchKernelGGL(k_film_encode, dim3(blocks_for(npix, 256)), dim3(256), 0, st, film, npix, width, thresholds, ldr,
                       rgbe, first_bad);
    return hipGetLastError();
}

hipError_t launch_film_rle(const uint8_t* rgbe, uint32_t width, uint32_t height, uint8_t* streams, uint32_t* lens,
                           hipStream_t st) {
    hipLaunchKernelGGL(k_film_rle, dim3(blocks_for(uint64_t(height) * 4u, 256)), dim3(256), 0, st, rgbe, width, height,
                       streams, lens);
    return hipGetLastError();
}

hipError_t set_turn_table(const TurnCS* table) {
    return hipMemcpyToSymbol(HIP_SYMBOL(g_turn_table), &table, sizeof(table));
}

hipError_t launch_accumulate(const DevJob* jobs, uint32_t job, uint64_t npix, float* film, hipStream_t st) {
    hipLaunchKernelGGL(k_accumulate, dim3(blocks_for(npix, 256)), dim3(256), 0, st, jobs, job, film);
    return hipGetLastError();
}

}  // namespace cr
