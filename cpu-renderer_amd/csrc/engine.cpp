// engine.cpp — job admission, stepping and completion of the wavefront engine.
#include "engine.h"

#include <algorithm>
#include <cstring>
#include <stdexcept>

namespace cr {

namespace {

constexpr uint32_t kMinPool = 1u << 12, kMaxPool = 1u << 25;
// A path finishes within 2*(bounces+2) <= 132 iterations of its start (one
// precision replay at most), so host steps without any retirement or issue
// while jobs are admitted mean the engine is stuck.
constexpr int kStallSteps = 96;

uint64_t pool_bytes_for(uint32_t paths) {
    // ray + state records, 4 float4 + 2 u32 of shadow records, 7 u32 queues.
    return uint64_t(paths) * (sizeof(PathRay) + sizeof(PathRec) + 4 * 16 + 2 * 4 + 7 * 4 + 2 * sizeof(Pending));
}

double seconds_between(Clock::time_point a, Clock::time_point b) {
    return std::chrono::duration<double>(b - a).count();
}

}  // namespace

struct Job {
    uint64_t ticket = 0;
    Request req;
    uint64_t items = 0, npix = 0, bytes = 0, item_base = 0;
    uint32_t seq = 0, slot = 0;
    uint64_t first_step = 0;
    uint64_t retired = 0;  // camera samples finished, as of the last step read back
    void* d_samples = nullptr;
    float* d_film = nullptr;
    DevLightEntry* d_lights = nullptr;
    float* h_film = nullptr;
    hipEvent_t done = nullptr;
    enum State { kPending, kRunning, kFinishing, kDone } state = kPending;
    int status = CRG_OK;
    std::string error;
    bool has_err = false;
    ErrorRec err{};
    crg_stats stats{};
    Clock::time_point t_submit, t_admit, t_done;
};

// ------------------------------------------------------------------ profiler
Profiler::~Profiler() {
    for (hipEvent_t e : ev) (void)hipEventDestroy(e);
}

void Profiler::begin(int slot, hipStream_t st) {
    if (!on) return;
    if (ev.empty()) {
        ev.resize(2 * kPairs);
        for (auto& e : ev) check(hipEventCreate(&e), "hipEventCreate");
    }
    if (used == kPairs) drain(st);
    stage_of[used] = slot;
    check(hipEventRecord(ev[2 * used], st), "hipEventRecord");
}

void Profiler::end(hipStream_t st) {
    if (!on) return;
    check(hipEventRecord(ev[2 * used + 1], st), "hipEventRecord");
    ++used;
}

void Profiler::drain(hipStream_t st) {
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

void Profiler::clear() {
    std::fill(std::begin(total_ms), std::end(total_ms), 0.0);
    std::fill(std::begin(launches), std::end(launches), 0ull);
}

const char* Profiler::name(int slot) { return slot < kStageCount ? kStageNames[slot] : "k_accumulate"; }

// ------------------------------------------------------------------ engine
Engine::Engine(int device, unsigned compute_units) : device_(device), cus_(compute_units) {
    check(hipSetDevice(device), "hipSetDevice");
    check(hipStreamCreateWithFlags(&st_, hipStreamNonBlocking), "hipStreamCreate");
    check(hipHostMalloc(reinterpret_cast<void**>(&h_jobs_), sizeof(DevJob) * kJobSlots, hipHostMallocDefault),
          "hipHostMalloc");
    for (Readback& rb : rb_) {
        check(hipHostMalloc(reinterpret_cast<void**>(&rb.ctl), sizeof(Control), hipHostMallocDefault), "hipHostMalloc");
        check(hipHostMalloc(reinterpret_cast<void**>(&rb.cnt), sizeof(JobCounters), hipHostMallocDefault),
              "hipHostMalloc");
        check(hipEventCreateWithFlags(&rb.ev, hipEventDisableTiming), "hipEventCreate");
    }
    hipMemPool_t mp = nullptr;
    if (hipDeviceGetDefaultMemPool(&mp, device) == hipSuccess && mp) {
        uint64_t keep = UINT64_MAX;  // keep freed request buffers cached for the next admissions
        (void)hipMemPoolSetAttribute(mp, hipMemPoolAttrReleaseThreshold, &keep);
    }
    size_t free_b = 0, total_b = 0;
    if (hipMemGetInfo(&free_b, &total_b) == hipSuccess && total_b)
        budget_ = std::min<uint64_t>(budget_, uint64_t(total_b) - std::min<uint64_t>(total_b / 16, 8ull << 30));
    // 32 M path slots (12.9 GB): longer persistent traversal runs and a third
    // of the iterations of a 4 M pool (config-2 render 289 -> 339 Msamples/s);
    // four 256-lane workgroups per CU (the traversal kernels' occupancy), two
    // iterations per graph launch.
    tune_ = crg_tuning{1u << 25, 256, 4, 2, 4};
    slot_job_.resize(kJobSlots);
    worker_ = std::thread([this] { run(); });
}

Engine::~Engine() {
    {
        std::lock_guard<std::mutex> lk(mu_);
        stop_ = true;
        for (auto& j : pending_) {
            j->status = CRG_EDEVICE;
            j->error = "context destroyed before the request was admitted";
            j->state = Job::kDone;
        }
        pending_.clear();
    }
    cv_work_.notify_all();
    cv_done_.notify_all();
    if (worker_.joinable()) worker_.join();
    (void)hipSetDevice(device_);
    (void)hipStreamSynchronize(st_);
    if (graph_) (void)hipGraphExecDestroy(graph_);
    for (Readback& rb : rb_) {
        (void)hipHostFree(rb.ctl);
        (void)hipHostFree(rb.cnt);
        (void)hipEventDestroy(rb.ev);
    }
    for (auto& kv : pinned_cache_) (void)hipHostFree(kv.second);
    (void)hipHostFree(h_jobs_);
    (void)hipStreamDestroy(st_);
}

uint64_t Engine::footprint(const Request& r) {
    const uint64_t npix = uint64_t(r.win[2]) * uint64_t(r.win[3]);
    return npix * uint64_t(r.options.samples) * sizeof(float4) + npix * 3 * sizeof(float) +
           std::max<size_t>(r.lights.entries.size(), 1) * sizeof(DevLightEntry);
}

uint64_t Engine::submit(Request r) {
    auto j = std::make_shared<Job>();
    j->npix = uint64_t(r.win[2]) * uint64_t(r.win[3]);
    j->items = j->npix * uint64_t(r.options.samples);
    j->bytes = footprint(r);
    j->req = std::move(r);
    j->t_submit = Clock::now();
    std::lock_guard<std::mutex> lk(mu_);
    if (j->bytes + pool_bytes_now() + resident_ > budget_) throw std::bad_alloc();
    j->ticket = next_ticket_++;
    tickets_[j->ticket] = j;
    pending_.push_back(j);
    cv_work_.notify_all();
    return j->ticket;
}

int Engine::wait(uint64_t ticket, double timeout_seconds, crg_stats* stats, std::string* err) {
    std::unique_lock<std::mutex> lk(mu_);
    auto it = tickets_.find(ticket);
    if (it == tickets_.end()) throw std::out_of_range("unknown request ticket");
    std::shared_ptr<Job> j = it->second;
    auto done = [&] { return j->state == Job::kDone; };
    if (timeout_seconds < 0) {
        cv_done_.wait(lk, done);
    } else if (!cv_done_.wait_for(lk, std::chrono::duration<double>(timeout_seconds), done)) {
        return CRG_EBUSY;
    }
    tickets_.erase(ticket);
    if (stats) *stats = j->stats;
    if (err) *err = j->error;
    return j->status;
}

void Engine::progress(uint64_t ticket, uint64_t* complete, uint64_t* total) {
    std::lock_guard<std::mutex> lk(mu_);
    auto it = tickets_.find(ticket);
    if (it == tickets_.end()) throw std::out_of_range("unknown request ticket");
    const Job& j = *it->second;
    *total = j.items;
    *complete = j.state == Job::kFinishing || j.state == Job::kDone ? j.items : j.retired;
}

void Engine::set_budget(uint64_t bytes) {
    std::lock_guard<std::mutex> lk(mu_);
    budget_ = bytes;
    cv_work_.notify_all();
}

void Engine::add_resident(int64_t delta) {
    std::lock_guard<std::mutex> lk(mu_);
    resident_ = uint64_t(int64_t(resident_) + delta);
}

crg_engine_info Engine::info() {
    std::lock_guard<std::mutex> lk(mu_);
    crg_engine_info i{};
    i.budget_bytes = budget_;
    i.resident_bytes = resident_;
    i.inflight_bytes = inflight_bytes_ + pool_bytes_now();
    i.inflight = uint32_t(running_.size() + finishing_.size());
    i.pending = uint32_t(pending_.size());
    i.completed = completed_;
    i.steps = steps_;
    return i;
}

crg_tuning Engine::tuning() {
    std::lock_guard<std::mutex> lk(mu_);
    return tune_;
}

void Engine::set_tuning(const crg_tuning& t) {
    if (t.pool_paths < kMinPool || t.pool_paths > kMaxPool) throw std::invalid_argument("pool_paths out of range");
    if (t.workgroup != 64 && t.workgroup != 128 && t.workgroup != 256)
        throw std::invalid_argument("workgroup must be 64, 128 or 256");
    if (t.groups_per_cu < 1 || t.groups_per_cu > 64) throw std::invalid_argument("groups_per_cu out of range");
    if (t.iterations_per_graph < 2 || t.iterations_per_graph > 64 || (t.iterations_per_graph & 1))
        throw std::invalid_argument("iterations_per_graph must be even and in [2, 64]");
    if (t.bvh_width != 4 && t.bvh_width != 8) throw std::invalid_argument("bvh_width must be 4 or 8");
    std::unique_lock<std::mutex> lk(mu_);
    cv_done_.wait(lk, [&] { return pending_.empty() && running_.empty() && finishing_.empty(); });
    if (std::memcmp(&t, &tune_, sizeof(t)) != 0) {
        tune_ = t;
        dirty_ = true;
    }
}

void Engine::set_counting(bool on) {
    std::unique_lock<std::mutex> lk(mu_);
    cv_done_.wait(lk, [&] { return pending_.empty() && running_.empty() && finishing_.empty(); });
    if (on != count_) {
        count_ = on;
        dirty_ = true;  // the graph holds the other traversal variants
    }
}

void Engine::set_profiling(bool on) {
    std::unique_lock<std::mutex> lk(mu_);
    cv_done_.wait(lk, [&] { return running_.empty() && finishing_.empty(); });
    prof_.on = on;
    if (on) prof_.clear();
}

void Engine::kernel_times(crg_kernel_time* out, int max, int* count) {
    std::lock_guard<std::mutex> lk(mu_);
    const int n = Profiler::kSlots;
    if (count) *count = n;
    for (int i = 0; i < n && i < max && out; ++i) {
        std::memset(&out[i], 0, sizeof(out[i]));
        std::strncpy(out[i].name, Profiler::name(i), sizeof(out[i].name) - 1);
        out[i].launches = prof_.launches[i];
        out[i].total_ms = prof_.total_ms[i];
    }
}

// Path pool bytes of the configuration the next request will run with.
uint64_t Engine::pool_bytes_now() const { return dirty_ ? pool_bytes_for(tune_.pool_paths) : pool_bytes_; }

float* Engine::pinned_get(size_t bytes) {
    auto it = pinned_cache_.lower_bound(bytes);
    if (it != pinned_cache_.end() && it->first <= bytes * 2) {
        float* p = it->second;
        pinned_cache_.erase(it);
        return p;
    }
    void* p = nullptr;
    check(hipHostMalloc(&p, std::max<size_t>(bytes, 64), hipHostMallocDefault), "hipHostMalloc");
    return static_cast<float*>(p);
}

void Engine::pinned_put(float* p, size_t bytes) {
    if (p) pinned_cache_.emplace(std::max<size_t>(bytes, 64), p);
}

// Pool buffers, control block, job table and the captured iteration graph.
void Engine::configure() {
    check(hipStreamSynchronize(st_), "configure");
    if (graph_) {
        check(hipGraphExecDestroy(graph_), "hipGraphExecDestroy");
        graph_ = nullptr;
    }
    const uint32_t n = tune_.pool_paths;
    const size_t f4 = size_t(n) * sizeof(float4), u32 = size_t(n) * 4;
    ray_.alloc(size_t(n) * sizeof(PathRay));
    rec_.alloc(size_t(n) * sizeof(PathRec));
    for (DevBuf* b : {&sh_o_, &sh_d_, &sh_lo_, &sh_c_}) b->alloc(f4);
    for (DevBuf* b : {&sh_path_, &sh_job_, &q0_, &q1_, &retire_, &rjob_, &xq_, &xsq_, &freeq_})
        b->alloc(u32);
    for (DevBuf* b : {&xpend_, &xspend_}) b->alloc(size_t(n) * sizeof(Pending));
    P_ = PathSoA{ray_.as<PathRay>(), rec_.as<PathRec>()};
    S_ = ShadowSoA{sh_path_.as<uint32_t>(), sh_job_.as<uint32_t>(), sh_o_.as<float4>(), sh_d_.as<float4>(),
                   sh_lo_.as<float4>(), sh_c_.as<float4>()};
    Q_ = Queues{{q0_.as<uint32_t>(), q1_.as<uint32_t>()}, retire_.as<uint32_t>(), rjob_.as<uint32_t>(),
                xq_.as<uint32_t>(),
                xsq_.as<uint32_t>(), freeq_.as<uint32_t>(), xpend_.as<Pending>(), xspend_.as<Pending>()};
    ctl_.alloc(sizeof(Control));
    cnt_.alloc(sizeof(JobCounters));
    jobs_.alloc(sizeof(DevJob) * kJobSlots);
    check(hipMemsetAsync(cnt_.p, 0, sizeof(JobCounters), st_), "hipMemsetAsync");
    check(hipMemsetAsync(ctl_.p, 0, sizeof(Control), st_), "hipMemsetAsync");
    masks_ = false;
    check(launch_pool_init(Q_.freeq, n, ctl_.as<Control>(), st_), "pool init");
    next_seq_ = seq_lo_ = 0;
    total_items_ = 0;
    err_read_ = 0;
    pool_bytes_ = pool_bytes_for(n);

    const LaunchShape shape{cus_ * tune_.groups_per_cu, tune_.workgroup, count_, int(tune_.bvh_width)};
    check(hipStreamSynchronize(st_), "configure");
    check(hipStreamBeginCapture(st_, hipStreamCaptureModeRelaxed), "hipStreamBeginCapture");
    hipError_t le = hipSuccess;
    for (uint32_t k = 0; k < tune_.iterations_per_graph; ++k)
        for (int stage = 0; stage < kStageCount && le == hipSuccess; ++stage)
            le = launch_stage(Stage(stage), P_, S_, jobs_.as<DevJob>(), Q_, int(k & 1), ctl_.as<Control>(),
                              cnt_.as<JobCounters>(), shape, st_);
    hipGraph_t g = nullptr;
    const hipError_t ce = hipStreamEndCapture(st_, &g);
    check(le, "graph capture launch");
    check(ce, "hipStreamEndCapture");
    const hipError_t ie = hipGraphInstantiate(&graph_, g, nullptr, nullptr, 0);
    (void)hipGraphDestroy(g);
    check(ie, "hipGraphInstantiate");
    configured_ = true;
    dirty_ = false;
}

void Engine::admit_locked() {
    bool any = false;
    while (!pending_.empty()) {
        std::shared_ptr<Job> jp = pending_.front();
        Job& j = *jp;
        if (running_.size() + finishing_.size() >= kJobSlots) break;
        const uint32_t seq = next_seq_, slot = seq % kJobSlots;
        if (slot_job_[slot] || (seq >= kJobSlots && seq_lo_ <= seq - kJobSlots)) break;
        if (resident_ + pool_bytes_ + inflight_bytes_ + j.bytes > budget_) break;
        pending_.pop_front();
        const size_t nl = std::max<size_t>(j.req.lights.entries.size(), 1);
        hipError_t e = hipMallocAsync(&j.d_samples, j.items * sizeof(float4), st_);
        if (e == hipSuccess) e = hipMallocAsync(reinterpret_cast<void**>(&j.d_film), j.npix * 3 * sizeof(float), st_);
        if (e == hipSuccess) e = hipMallocAsync(reinterpret_cast<void**>(&j.d_lights), nl * sizeof(DevLightEntry), st_);
        if (e == hipSuccess) e = hipEventCreateWithFlags(&j.done, hipEventDisableTiming);
        if (e != hipSuccess) {
            (void)hipGetLastError();
            release_job_memory(j);
            fail_locked(j, e == hipErrorOutOfMemory ? CRG_ENOMEM : CRG_EDEVICE,
                        e == hipErrorOutOfMemory ? "out of memory" : std::string("admission: ") + hipGetErrorString(e));
            continue;
        }
        j.h_film = pinned_get(j.npix * 3 * sizeof(float));
        const crg_options& o = j.req.options;
        DevJob& d = h_jobs_[slot];
        d = DevJob{};
        d.scene = j.req.scene->host;
        d.env = j.req.env->host;
        d.lights = DevLights{j.d_lights, uint32_t(j.req.lights.entries.size()), j.req.lights.env_probability};
        d.cam = j.req.camera;
        d.width = o.width;
        d.height = o.height;
        d.spp = o.samples;
        d.bounces = o.bounces;
        d.seed = o.seed;
        d.nee = o.next_event_estimation;
        d.rr = o.russian_roulette;
        d.samples = static_cast<float4*>(j.d_samples);
        d.item_base = total_items_;
        d.items = j.items;
        d.win_x = j.req.win[0];
        d.win_y = j.req.win[1];
        d.win_w = j.req.win[2];
        d.win_h = j.req.win[3];
        if (!j.req.lights.entries.empty())
            check(hipMemcpyAsync(j.d_lights, j.req.lights.entries.data(), j.req.lights.entries.size() * sizeof(DevLightEntry),
                                 hipMemcpyHostToDevice, st_),
                  "light table upload");
        check(hipMemcpyAsync(jobs_.as<DevJob>() + slot, &d, sizeof(DevJob), hipMemcpyHostToDevice, st_), "job upload");
        j.item_base = total_items_;
        total_items_ += j.items;
        j.seq = seq;
        j.slot = slot;
        ++next_seq_;
        check(launch_admit(ctl_.as<Control>(), cnt_.as<JobCounters>(), total_items_, seq_lo_, next_seq_, slot, st_),
              "admit");
        j.first_step = steps_;
        j.state = Job::kRunning;
        j.t_admit = Clock::now();
        slot_job_[slot] = jp;
        running_.push_back(jp);
        inflight_bytes_ += j.bytes;
        any = true;
    }
    if (any) {
        update_masks_locked();
        // Idle slots start the new items right away (same reset + refill the
        // graph ends with; both are idempotent at a step boundary).
        const LaunchShape shape{cus_ * tune_.groups_per_cu, tune_.workgroup, count_, int(tune_.bvh_width)};
        for (Stage s : {kStageReset, kStageRefill})
            check(launch_stage(s, P_, S_, jobs_.as<DevJob>(), Q_, 1, ctl_.as<Control>(), cnt_.as<JobCounters>(), shape,
                               st_),
                  "refill");
    }
}

void Engine::update_masks_locked() {
    bool m = false;
    for (auto& jp : running_) m = m || jp->req.scene->masks;
    if (m == masks_) return;
    check(launch_set_masks(ctl_.as<Control>(), m ? 1u : 0u, st_), "traversal variant");
    masks_ = m;
}

void Engine::launch_step(Readback& rb) {
    if (prof_.on) {
        const LaunchShape shape{cus_ * tune_.groups_per_cu, tune_.workgroup, count_, int(tune_.bvh_width)};
        for (uint32_t k = 0; k < tune_.iterations_per_graph; ++k)
            for (int stage = 0; stage < kStageCount; ++stage) {
                prof_.begin(stage, st_);
                check(launch_stage(Stage(stage), P_, S_, jobs_.as<DevJob>(), Q_, int(k & 1), ctl_.as<Control>(),
                                   cnt_.as<JobCounters>(), shape, st_),
                      "iteration launch");
                prof_.end(st_);
            }
    } else {
        check(hipGraphLaunch(graph_, st_), "hipGraphLaunch");
    }
    check(hipMemcpyAsync(rb.ctl, ctl_.p, sizeof(Control), hipMemcpyDeviceToHost, st_), "control read");
    check(hipMemcpyAsync(rb.cnt, cnt_.p, sizeof(JobCounters), hipMemcpyDeviceToHost, st_), "counter read");
    check(hipEventRecord(rb.ev, st_), "hipEventRecord");
    if (prof_.on) prof_.drain(st_);
}

void Engine::process_locked(const Readback& rb) {
    const Control& c = *rb.ctl;
    const JobCounters& n = *rb.cnt;
    if (c.err_count != err_read_) {
        uint32_t from = err_read_;
        if (c.err_count - from > uint32_t(kMaxErrors)) from = c.err_count - uint32_t(kMaxErrors);
        for (uint32_t i = from; i != c.err_count; ++i) {
            const ErrorRec& e = c.errors[i % uint32_t(kMaxErrors)];
            const std::shared_ptr<Job>& jp = slot_job_[e.job % kJobSlots];
            if (!jp || jp->state != Job::kRunning || rb.step < jp->first_step) continue;
            if (!jp->has_err || e.key < jp->err.key) {
                jp->err = e;
                jp->has_err = true;
            }
        }
        err_read_ = c.err_count;
    }
    uint64_t progress = c.issued;
    std::vector<std::shared_ptr<Job>> still;
    still.reserve(running_.size());
    for (auto& jp : running_) {
        Job& j = *jp;
        if (rb.step >= j.first_step) {
            progress += n.retired[j.slot];
            j.retired = std::min<uint64_t>(n.retired[j.slot], j.items);
            if (n.retired[j.slot] >= j.items) {
                j.stats.iterations = int32_t((rb.step + 1 - j.first_step) * tune_.iterations_per_graph);
                complete_locked(jp, n);
                continue;
            }
        }
        still.push_back(jp);
    }
    running_.swap(still);
    update_masks_locked();
    // Oldest job whose items are not all issued yet (lookups never go below it).
    uint32_t lo = next_seq_;
    for (auto& jp : running_)
        if (jp->item_base + jp->items > c.issued) lo = std::min(lo, jp->seq);
    seq_lo_ = std::max(seq_lo_, lo);
    if (running_.empty() || progress != last_progress_) {
        stall_steps_ = 0;
    } else if (++stall_steps_ > kStallSteps) {
        stall_steps_ = 0;
        throw std::logic_error("render did not converge");
    }
    last_progress_ = progress;
}

void Engine::complete_locked(const std::shared_ptr<Job>& jp, const JobCounters& n) {
    Job& j = *jp;
    const uint32_t s = j.slot;
    j.stats.camera_samples = j.items;
    j.stats.boundary_retries = n.replays[s];
    j.stats.extend_rays = n.extend_rays[s];
    j.stats.shadow_rays = n.shadow_rays[s];
    j.stats.extend_nodes = n.extend_nodes[s];
    j.stats.extend_tris = n.extend_tris[s];
    j.stats.shadow_nodes = n.shadow_nodes[s];
    j.stats.shadow_tris = n.shadow_tris[s];
    j.stats.exact_fallbacks = n.fallbacks[s];
    j.stats.exact_rays = n.exact_rays[s];
    prof_.begin(kStageCount, st_);
    check(launch_accumulate(jobs_.as<DevJob>(), s, j.npix, j.d_film, st_), "accumulate launch");
    prof_.end(st_);
    check(hipMemcpyAsync(j.h_film, j.d_film, j.npix * 3 * sizeof(float), hipMemcpyDeviceToHost, st_), "film read");
    check(hipEventRecord(j.done, st_), "hipEventRecord");
    j.state = Job::kFinishing;
    finishing_.push_back(jp);
}

void Engine::finalize_ready_locked(bool block) {
    std::vector<std::shared_ptr<Job>> keep;
    for (auto& jp : finishing_) {
        Job& j = *jp;
        if (block) {
            check(hipEventSynchronize(j.done), "film read");
        } else {
            const hipError_t q = hipEventQuery(j.done);
            if (q == hipErrorNotReady) {
                keep.push_back(jp);
                continue;
            }
            check(q, "film read");
        }
        j.t_done = Clock::now();
        if (j.status == CRG_OK) {
            std::memcpy(j.req.film, j.h_film, j.npix * 3 * sizeof(float));
            if (j.req.samples) {  // each sample's radiance, for callers comparing samples one by one
                std::vector<float4> v(j.items);
                check(hipMemcpy(v.data(), j.d_samples, j.items * sizeof(float4), hipMemcpyDeviceToHost), "samples read");
                for (uint64_t i = 0; i < j.items; ++i) {
                    j.req.samples[3 * i] = v[i].x;
                    j.req.samples[3 * i + 1] = v[i].y;
                    j.req.samples[3 * i + 2] = v[i].z;
                }
            }
            if (j.has_err) {
                j.status = CRG_EPATH;
                j.error = format_path_error(j.err, j.req.options.width, j.req.options.samples);
            }
        }
        j.stats.seconds = seconds_between(j.t_admit, j.t_done);
        j.stats.queue_seconds = seconds_between(j.t_submit, j.t_admit);
        j.stats.latency_seconds = seconds_between(j.t_submit, j.t_done);
        j.stats.workers = 0;
        release_job_memory(j);
        slot_job_[j.slot].reset();
        inflight_bytes_ -= j.bytes;
        ++completed_;
        j.req.scene.reset();
        j.req.env.reset();
        j.state = Job::kDone;
    }
    const bool changed = keep.size() != finishing_.size();
    finishing_.swap(keep);
    if (changed) {
        cv_done_.notify_all();
        cv_work_.notify_all();
    }
}

void Engine::fail_locked(Job& j, int status, const std::string& msg) {
    j.status = status;
    j.error = msg;
    j.t_done = Clock::now();
    j.stats.latency_seconds = seconds_between(j.t_submit, j.t_done);
    j.req.scene.reset();
    j.req.env.reset();
    j.state = Job::kDone;
    cv_done_.notify_all();
}

void Engine::release_job_memory(Job& j) {
    if (j.d_samples) (void)hipFreeAsync(j.d_samples, st_);
    if (j.d_film) (void)hipFreeAsync(j.d_film, st_);
    if (j.d_lights) (void)hipFreeAsync(j.d_lights, st_);
    if (j.done) (void)hipEventDestroy(j.done);
    pinned_put(j.h_film, j.npix * 3 * sizeof(float));
    j.d_samples = nullptr;
    j.d_film = nullptr;
    j.d_lights = nullptr;
    j.done = nullptr;
    j.h_film = nullptr;
}

void Engine::run() {
    (void)hipSetDevice(device_);
    std::unique_lock<std::mutex> lk(mu_);
    std::deque<int> inflight;  // readback slots of launched, unprocessed steps
    for (;;) {
        if (running_.empty() && finishing_.empty() && inflight.empty()) {
            cv_work_.wait(lk, [&] { return stop_ || !pending_.empty(); });
            if (pending_.empty()) break;  // stop requested and nothing left
        }
        try {
            if (dirty_ && running_.empty() && finishing_.empty() && inflight.empty()) configure();
            admit_locked();
            const bool have = !running_.empty();
            if (have) {
                const int b = int(steps_ & 1);
                Readback& rb = rb_[b];
                rb.step = steps_;
                lk.unlock();
                launch_step(rb);
                lk.lock();
                ++steps_;
                inflight.push_back(b);
            }
            if (!inflight.empty() && (inflight.size() >= 2 || !have)) {
                const int b = inflight.front();
                inflight.pop_front();
                lk.unlock();
                const hipError_t e = hipEventSynchronize(rb_[b].ev);
                lk.lock();
                check(e, "wavefront step");
                process_locked(rb_[b]);
            }
            finalize_ready_locked(running_.empty() && inflight.empty());
            if (!have && inflight.empty() && finishing_.empty() && !pending_.empty() && running_.empty()) {
                // Nothing in flight and the head of the queue does not fit:
                // it never will (budget shrank below it).
                Job& j = *pending_.front();
                if (resident_ + pool_bytes_now() + j.bytes > budget_) {
                    std::shared_ptr<Job> keep = pending_.front();
                    pending_.pop_front();
                    fail_locked(j, CRG_ENOMEM, "out of memory");
                } else {
                    cv_work_.wait_for(lk, std::chrono::milliseconds(5));
                }
            }
        } catch (const std::exception& ex) {
            // Fail every request; the pool is rebuilt before the next one (a
            // device that stays broken fails that one too).
            const std::string msg = ex.what();
            const int code = dynamic_cast<const std::bad_alloc*>(&ex)   ? CRG_ENOMEM
                             : dynamic_cast<const std::logic_error*>(&ex) ? CRG_EINVAL
                                                                           : CRG_EDEVICE;
            for (auto* list : {&running_, &finishing_}) {
                for (auto& jp : *list) {
                    release_job_memory(*jp);
                    fail_locked(*jp, code, msg);
                }
                list->clear();
            }
            for (auto& jp : pending_) fail_locked(*jp, code, msg);
            pending_.clear();
            inflight.clear();
            for (auto& sj : slot_job_) sj.reset();
            inflight_bytes_ = 0;
            dirty_ = true;
            configured_ = false;
            (void)hipGetLastError();
        }
    }
}

}  // namespace cr
