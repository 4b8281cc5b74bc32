// engine.h — the wavefront engine shared by plain, batched and served renders.
//
// Every render request becomes a job: its camera samples are a contiguous
// range of one global work-item stream, and one pool of path slots runs the
// items of all admitted jobs.  A worker thread owns the device stream: it
// admits jobs into the device job table (continuous batching: a job admitted
// while others run starts in the slots their finished paths free), launches
// the wavefront iterations as one captured hipGraph per host step, reads back
// the control block and per-job counters, and completes jobs whose samples
// have all retired (per-pixel accumulation, film readback).  Results depend
// only on (scene, camera, options) — never on slots, batch composition or the
// launch configuration — so a request renders bit-identically alone or in a
// mix (the reference's fixed per-pixel sample order, src/render.cpp:384-415).
#pragma once
#include <hip/hip_runtime.h>

#include <chrono>
#include <condition_variable>
#include <cstdint>
#include <deque>
#include <map>
#include <memory>
#include <mutex>
#include <string>
#include <thread>
#include <vector>

#include "../../include/crgpu.h"
#include "device_buf.h"
#include "host_scene.h"
#include "kernels.h"

namespace cr {

using Clock = std::chrono::steady_clock;

struct SceneRes {
    DevBuf nodes4, nodes8, prims, prim_ng, aux, materials, light_rank, punctual, uvs, tangents, texels, images, textures, desc;
    DevScene host{};
    SceneLights lights;
    std::vector<uint32_t> tri_to_prim;
    std::vector<uint32_t> prim_to_tri;
    uint64_t triangles = 0, bytes = 0, binary_nodes = 0;
    uint64_t texel_count = 0, image_count = 0, texture_count = 0, material_count = 0;  // device array lengths
    bool masks = false;  // some primitive has an alpha-mask texture
};

struct EnvRes {
    DevBuf texels, pdf, threshold, guide, rowcos, desc;
    DevEnv host{};
    double integral = 0;
    uint64_t bytes = 0;
};

// Event-pair timing of kernel launches on the engine stream (profiling mode
// launches the iteration kernels one by one instead of as a graph).
struct Profiler {
    static constexpr int kSlots = kStageCount + 1;  // + accumulate
    static constexpr int kPairs = 256;
    bool on = false;
    std::vector<hipEvent_t> ev;
    int stage_of[kPairs] = {};
    int used = 0;
    double total_ms[kSlots] = {};
    uint64_t launches[kSlots] = {};

    ~Profiler();
    void begin(int slot, hipStream_t st);
    void end(hipStream_t st);
    void drain(hipStream_t st);
    void clear();
    static const char* name(int slot);
};

// A validated request, prepared on the submitting thread.
struct Request {
    std::shared_ptr<SceneRes> scene;
    std::shared_ptr<EnvRes> env;
    crg_options options{};
    DevCamera camera{};
    LightTable lights;
    float* film = nullptr;     // win_w * win_h * 3 floats
    int32_t win[4] = {0, 0, 0, 0};  // x, y, w, h of the frame window rendered (the whole frame by default)
    float* samples = nullptr;  // optional: spp * win_w * win_h * 3 floats, each sample's radiance
};

struct Job;

class Engine {
public:
    Engine(int device, unsigned compute_units);
    ~Engine();
    Engine(const Engine&) = delete;
    Engine& operator=(const Engine&) = delete;

    // Queue a request; throws std::bad_alloc if it can never fit the budget.
    uint64_t submit(Request r);
    // CRG_EBUSY on timeout, else the request's status (error text in *err).
    int wait(uint64_t ticket, double timeout_seconds, crg_stats* stats, std::string* err);
    // Camera samples of a submitted request finished so far (as of the last
    // step read back), and its total.
    void progress(uint64_t ticket, uint64_t* complete, uint64_t* total);

    void set_budget(uint64_t bytes);
    void add_resident(int64_t delta_bytes);
    crg_engine_info info();
    crg_tuning tuning();
    void set_tuning(const crg_tuning& t);  // waits until idle
    void set_profiling(bool on);
    // Traversal kernels with per-query work counters (node visits, triangle
    // tests) for the requests that follow; waits until idle.
    void set_counting(bool on);
    void kernel_times(crg_kernel_time* out, int max, int* count);

    static uint64_t footprint(const Request& r);

private:
    struct Readback {
        Control* ctl = nullptr;       // pinned
        JobCounters* cnt = nullptr;   // pinned
        hipEvent_t ev = nullptr;
        uint64_t step = 0;
    };

    void run();
    void configure();
    void admit_locked();
    void update_masks_locked();  // traversal variant for the live jobs' scenes
    void launch_step(Readback& rb);
    void process_locked(const Readback& rb);
    void complete_locked(const std::shared_ptr<Job>& j, const JobCounters& c);
    void finalize_ready_locked(bool block);
    void fail_locked(Job& j, int status, const std::string& msg);
    void release_job_memory(Job& j);
    uint64_t pool_bytes_now() const;
    float* pinned_get(size_t bytes);
    void pinned_put(float* p, size_t bytes);

    int device_;
    unsigned cus_;
    hipStream_t st_ = nullptr;
    std::thread worker_;
    std::mutex mu_;
    std::condition_variable cv_work_, cv_done_;
    bool stop_ = false;

    // configuration
    crg_tuning tune_{};
    bool dirty_ = true;
    bool configured_ = false;
    bool masks_ = false;  // ctl->masks as last set
    bool count_ = false;  // work-counting traversal variants
    hipGraphExec_t graph_ = nullptr;
    Profiler prof_;

    // device state
    DevBuf ray_, rec_;
    DevBuf sh_path_, sh_job_, sh_o_, sh_d_, sh_lo_, sh_c_;
    DevBuf q0_, q1_, retire_, rjob_, xq_, xsq_, freeq_, xpend_, xspend_;
    DevBuf ctl_, cnt_, jobs_;
    PathSoA P_{};
    ShadowSoA S_{};
    Queues Q_{};
    DevJob* h_jobs_ = nullptr;  // pinned mirror of the device job table
    Readback rb_[2];
    uint64_t pool_bytes_ = 0;

    // host bookkeeping
    uint64_t next_ticket_ = 1;
    uint32_t next_seq_ = 0, seq_lo_ = 0;
    uint64_t total_items_ = 0;
    uint64_t steps_ = 0;
    uint32_t err_read_ = 0;
    uint64_t budget_ = 200ull << 30;
    uint64_t resident_ = 0, inflight_bytes_ = 0, completed_ = 0;
    uint64_t last_progress_ = 0;
    int stall_steps_ = 0;
    std::deque<std::shared_ptr<Job>> pending_;
    std::vector<std::shared_ptr<Job>> running_, finishing_;
    std::vector<std::shared_ptr<Job>> slot_job_;  // kJobSlots
    std::map<uint64_t, std::shared_ptr<Job>> tickets_;
    std::multimap<size_t, float*> pinned_cache_;
};

}  // namespace cr
