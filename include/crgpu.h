/*
 * crgpu.h — C ABI of the MI355X (gfx950) wavefront path tracer.
 *
 * This is the drop-in boundary for the reference renderer's hot path:
 *   RenderResult cr::render(const Scene&, const Bvh&, const Camera&,
 *                           const Environment&, const RenderOptions&, Progress)
 *                                                   (reference src/render.h:53)
 *   bool Bvh::intersect(const Scene&, const Ray&, Hit&, bool any, uint32_t ignore)
 *                                                   (reference src/scene.h:126-127)
 *   void Bvh::build(const Scene&)                   (reference src/scene.h:123)
 *
 * Every entry point takes plain pointers and sizes, never throws, and returns
 * an int status (CRG_OK or one of the CRG_E* codes).  The text of the most
 * recent failure on a context is returned by crg_last_error(), in the same
 * wording the reference throws (e.g. "Pixel (x,y), sample s: <cause>",
 * reference src/render.cpp:437-438).
 *
 * Descriptor structs below mirror the reference scene model after import
 * (reference src/scene.h:10-117, src/render.h:8-39) as flat arrays.  The same
 * descriptors are consumed by the CPU oracle (oracle/) so that every parity
 * test feeds identical inputs to both sides.
 */
#ifndef CRGPU_H
#define CRGPU_H

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

/* ---- status codes ------------------------------------------------------ */
#define CRG_OK        0
#define CRG_EINVAL    1  /* invalid argument / option validation failure    */
#define CRG_ENOMEM    2  /* host or device allocation failed (bad_alloc)     */
#define CRG_EPATH     3  /* non-finite/negative transport or replay failure  */
#define CRG_EDEVICE   4  /* HIP runtime failure or no usable gfx950 device   */
#define CRG_ENOTFOUND 5  /* unknown scene / environment / request handle     */
#define CRG_EBUSY     6  /* request still in flight (crg_wait timed out)    */

/* ---- scene model ------------------------------------------------------- */
/* Texture binding: texture = -1 means "no map" (reference src/scene.h:49-53). */
typedef struct crg_texture_slot {
    int32_t texture, texcoord;
    float offset[2], scale[2], rotation;
} crg_texture_slot;

/* Material factors (reference src/scene.h:58-70).  alpha_mode: 0 OPAQUE, 1 MASK. */
typedef struct crg_material {
    float base_color[4];
    float metallic, roughness;
    float emission[3];
    float transmission, ior;
    float thickness, attenuation_distance;
    float attenuation_color[3];
    int32_t double_sided;
    int32_t alpha_mode;
    float alpha_cutoff, normal_scale;
    crg_texture_slot base_color_tex, metallic_roughness_tex, normal_tex, emissive_tex, transmission_tex;
} crg_material;

/* Punctual light (reference src/scene.h:79-86). type: 0 point, 1 directional, 2 spot. */
typedef struct crg_light {
    int32_t type;
    float position[3], direction[3], color[3];
    float intensity, range, inner_cone, outer_cone;
} crg_light;

/* Decoded image: rgba is width*height*4 floats (reference src/scene.h:47). */
typedef struct crg_image {
    int32_t width, height;
    const float* rgba;
} crg_image;

/* Sampler binding of an image. wrap: 0 repeat, 1 clamp, 2 mirror (reference src/scene.h:44-48). */
typedef struct crg_texture {
    int32_t image, wrap_s, wrap_t, linear;
} crg_texture;

/*
 * Triangle soup in world space, before Scene::finalize (reference
 * src/scene.h:24-34).  positions: 9 floats per triangle (v0,v1,v2).
 * normals: 9 floats per triangle or NULL (zero normals are repaired by
 * finalize).  tangents: 12 floats per triangle (xyzw per vertex) or NULL for
 * (0,0,0,1).  uvs: 12 floats per triangle laid out [vertex][set][uv] or NULL.
 */
typedef struct crg_scene_desc {
    uint64_t triangle_count;
    const float* positions;
    const float* normals;
    const float* tangents;
    const float* uvs;
    const uint32_t* material_ids;
    uint32_t material_count;
    const crg_material* materials;
    uint32_t light_count;
    const crg_light* lights;
    uint32_t image_count;
    const crg_image* images;
    uint32_t texture_count;
    const crg_texture* textures;
} crg_scene_desc;

/* Perspective camera (reference src/scene.h:71-78). */
typedef struct crg_camera {
    float position[3], forward[3], right[3], up[3];
    float yfov, aspect, near_clip, far_clip;
} crg_camera;

/*
 * Environment (reference src/render.h:8-28).  rgba == NULL selects the
 * constant environment `constant`; otherwise a width*height lat-long image
 * with rotation (degrees) and intensity.
 */
typedef struct crg_env_desc {
    float constant[3];
    int32_t width, height;
    const float* rgba;
    float rotation_degrees, intensity;
} crg_env_desc;

/* Render options (reference src/render.h:31-39). threads is ignored on the GPU. */
typedef struct crg_options {
    int32_t width, height, samples, bounces, threads;
    int32_t next_event_estimation, russian_roulette;
    uint64_t seed;
} crg_options;

/* Result statistics (reference src/render.h:41-48 plus device counters). */
typedef struct crg_stats {
    double seconds;            /* render time, excluding upload and BVH build  */
    uint64_t camera_samples;
    uint64_t boundary_retries; /* precision replays (reference render.cpp:409) */
    uint64_t extend_rays;      /* closest-hit queries, replays included        */
    uint64_t shadow_rays;      /* any-hit queries                              */
    int32_t workers;           /* 0 = device                                   */
    int32_t iterations;        /* wavefront iterations executed                */
    /* Work counters, filled only while crg_set_work_counters is on (0
     * otherwise).  They depend on how queries share waves, so a request's
     * values differ slightly between runs alone and in a batch. */
    uint64_t extend_nodes;     /* wide nodes visited by closest-hit queries    */
    uint64_t extend_tris;      /* triangles tested by closest-hit queries      */
    uint64_t shadow_nodes;     /* wide nodes visited by any-hit queries        */
    uint64_t shadow_tris;      /* triangles tested by any-hit queries          */
    uint64_t exact_fallbacks;  /* triangle tests decided in emulated binary64  */
    uint64_t exact_rays;       /* queries rerun by the exact traversal         */
    double queue_seconds;      /* submission -> admission (batched / served)   */
    double latency_seconds;    /* submission -> film written (time to image)   */
} crg_stats;

/*
 * Ray query for the traversal parity harness (reference src/scene.h:8-14):
 * the ray is origin + origin_low + t*direction on [tmin, tmax]; `ignore` is
 * the excluded source triangle (0xffffffff = none, reference src/scene.h:124-127).
 */
typedef struct crg_ray {
    float origin[3], direction[3];
    float tmin, tmax;
    float origin_low[3];
    uint32_t ignore;
} crg_ray;

/* Closest/any-hit result.  t,u,v are the binary64 values the reference's
 * Hit holds (reference src/scene.h:35-37); the device computes them with an
 * exact integer emulation of binary64 and never executes FP64 instructions. */
typedef struct crg_hit {
    uint32_t hit, triangle;
    double t, u, v;
} crg_hit;

/* Scene metadata returned by upload. */
typedef struct crg_scene_info {
    uint64_t triangles;        /* after finalize (degenerates removed) */
    uint64_t bvh_nodes;        /* binary SAH nodes built on the host   */
    uint64_t wide_nodes;       /* 4-wide device nodes (64 B, quantized boxes); the
                                  8-wide collapse has about 0.7 as many */
    uint64_t device_bytes;     /* resident HBM footprint               */
    double build_seconds;      /* finalize + SAH build + flatten       */
    double upload_seconds;
} crg_scene_info;

typedef struct crg_ctx crg_ctx;

/* ---- context ----------------------------------------------------------- */
int crg_context_create(int device, crg_ctx** out);
void crg_context_destroy(crg_ctx* ctx);
const char* crg_last_error(const crg_ctx* ctx);
/* Library build identity ("gfx950 ..."), usable without a device. */
const char* crg_version(void);

/* ---- resident scenes / environments ------------------------------------ */
/* Scene::finalize + Bvh::build on the host, then upload (reference
 * src/geometry.cpp:296-397, :463-550).  Returns a resident handle. */
int crg_scene_upload(crg_ctx* ctx, const crg_scene_desc* desc, uint64_t* scene_id, crg_scene_info* info);
/*
 * Scene cache file.  No counterpart in the reference (its scenes are rebuilt
 * from glTF on every run); for servers that keep many scenes: crg_scene_save
 * writes the device-format arrays of a resident scene (wide BVH, primitive and
 * shading records, materials, textures, light data) with a checksum;
 * crg_scene_load makes such a file resident again without Scene::finalize or
 * the BVH build (info->build_seconds then counts the file read).  Renders of a
 * loaded scene are bit-identical to renders of the scene it was saved from.
 * A file is only valid for builds with the same record layouts.
 */
int crg_scene_save(crg_ctx* ctx, uint64_t scene_id, const char* path);
/* Loading checks every index in the file (node children, leaf ranges,
 * triangle ids, materials, textures, images, texel ranges, lights, the
 * leaf-order maps) before anything is uploaded, and refuses a file that fails
 * with a message naming what is wrong. */
int crg_scene_load(crg_ctx* ctx, const char* path, uint64_t* scene_id, crg_scene_info* info);
/* Scene::finalize + Bvh::build + flattening on the host (reference
 * src/geometry.cpp:296-397, :463-550) written straight to a cache file, with
 * no device: scenes can be prepared offline and made resident later by
 * crg_scene_load.  info->device_bytes is 0.  Without a context, failure text
 * goes to message (capacity bytes, always terminated; may be NULL). */
int crg_scene_write_cache(const crg_scene_desc* desc, const char* path, crg_scene_info* info, char* message,
                          size_t capacity);
/* The checks crg_scene_load makes (format, checksum, record layouts, every
 * index), without a device: CRG_OK, or CRG_EINVAL with the reason in message. */
int crg_scene_check(const char* path, char* message, size_t capacity);
int crg_scene_release(crg_ctx* ctx, uint64_t scene_id);
/* Environment construction (reference src/render.cpp:12-42). */
int crg_env_upload(crg_ctx* ctx, const crg_env_desc* desc, uint64_t* env_id, uint64_t* device_bytes);
int crg_env_release(crg_ctx* ctx, uint64_t env_id);

/* ---- hot path ---------------------------------------------------------- */
/* render(): film_rgb receives width*height*3 linear floats
 * (reference src/render.cpp:352-456). */
int crg_render(crg_ctx* ctx, uint64_t scene_id, uint64_t env_id, const crg_camera* camera,
               const crg_options* options, float* film_rgb, crg_stats* stats);

/* render() over a window of the frame: the pixels x0 <= x < x0+width,
 * y0 <= y < y0+height of the options' width x height image, each with its
 * frame index and hence exactly the sample stream, radiance and film value it
 * has in the full render (reference src/render.cpp:384-415).  film_rgb
 * receives width*height*3 floats; samples_rgb (may be NULL) receives every
 * sample's radiance, spp*width*height*3 floats ordered [sample][pixel].  No
 * counterpart in the reference: crops of large frames at high sample counts
 * (converged references, quality measurement) and sample-by-sample comparison. */
int crg_render_window(crg_ctx* ctx, uint64_t scene_id, uint64_t env_id, const crg_camera* camera,
                      const crg_options* options, int32_t x0, int32_t y0, int32_t width, int32_t height,
                      float* film_rgb, float* samples_rgb, crg_stats* stats);

/* Bvh::intersect over a host ray array (reference src/geometry.cpp:554-639).
 * any != 0 selects visibility queries (reference src/geometry.cpp:608). */
int crg_trace_rays(crg_ctx* ctx, uint64_t scene_id, const crg_ray* rays, uint64_t count,
                   crg_hit* hits, int any);

/* crg_trace_rays without any device allocation.  No counterpart in the
 * reference: for callers that cannot allocate device memory (code under HIP
 * graph capture, tight serving loops, harnesses that forbid hipMalloc).  Its
 * staging (control block, rays, hits) lives in a fixed 5.3 MB scratch array
 * that each device holds for the life of the process; at most
 * CRG_NOALLOC_MAX_RAYS rays per call (more: CRG_EINVAL, nothing launched).
 * Results equal crg_trace_rays's.  Calls on one context are serialized;
 * calls from two contexts on the same device share the array and race. */
#define CRG_NOALLOC_MAX_RAYS 65536u
int crg_trace_rays_noalloc(crg_ctx* ctx, uint64_t scene_id, const crg_ray* rays, uint64_t count,
                           crg_hit* hits, int any);

/* ---- batched and served forms ------------------------------------------
 * Requests over resident scenes share one wavefront path pool: their camera
 * samples are work items of one global stream, so bounces of renders of
 * different scenes and sizes run in the same kernel launches, and a request
 * admitted while others are in flight starts in the slots that finished paths
 * free up (continuous batching).  Each request's film is bit-identical to its
 * plain crg_render result.  The reference renders one call at a time
 * (reference src/render.cpp:352-456); these entry points have no counterpart
 * there and exist for callers that render many requests.
 */
typedef struct crg_request {
    uint64_t scene;        /* resident scene handle                            */
    uint64_t env;          /* resident environment handle                      */
    crg_camera camera;
    crg_options options;
    float* film_rgb;       /* width*height*3 floats, owned by the caller until
                              the request completes                            */
} crg_request;

/* Render `count` requests together; returns when all are done.  status[i] and
 * stats[i] (either may be NULL) receive each request's own outcome; the return
 * value is CRG_OK when every request succeeded, else the first failing
 * request's status (crg_last_error names it). */
int crg_render_batch(crg_ctx* ctx, const crg_request* requests, uint32_t count, int* status, crg_stats* stats);

/* Asynchronous submission (serving).  The request is validated now and
 * admitted as soon as its device footprint (sample buffer, film, light table)
 * fits the memory budget next to the resident scenes and the requests in
 * flight; a request that could never fit fails with CRG_ENOMEM. */
int crg_submit(crg_ctx* ctx, const crg_request* request, uint64_t* ticket);
/* Wait up to timeout_seconds (< 0: forever) for a submitted request.  Returns
 * CRG_EBUSY if it is still in flight, else the request's own status (its film
 * is written on CRG_OK); the ticket is consumed unless CRG_EBUSY. */
int crg_wait(crg_ctx* ctx, uint64_t ticket, double timeout_seconds, crg_stats* stats);

/* Camera samples of a submitted (not yet waited-for) request finished so far
 * and its total, as of the engine's last step: the reference's Progress
 * callback reports the same pair (reference src/render.cpp:419-427). */
int crg_request_progress(crg_ctx* ctx, uint64_t ticket, uint64_t* complete, uint64_t* total);

/* Device memory budget for admission (default 200 GiB, capped by the device). */
int crg_set_memory_budget(crg_ctx* ctx, uint64_t bytes);

typedef struct crg_engine_info {
    uint64_t budget_bytes;     /* admission budget                              */
    uint64_t resident_bytes;   /* resident scenes + environments                */
    uint64_t inflight_bytes;   /* path pool + admitted requests                 */
    uint32_t inflight;         /* admitted requests not yet complete            */
    uint32_t pending;          /* submitted requests waiting for admission      */
    uint64_t completed;        /* requests finished since context creation      */
    uint64_t steps;            /* host steps (graph launches)                   */
} crg_engine_info;
int crg_engine_info_get(crg_ctx* ctx, crg_engine_info* info);

/* ---- launch configuration and autotuning --------------------------------
 * Results never depend on these choices (every path's result is independent of
 * the slot, wave and iteration that ran it); only speed does. */
typedef struct crg_tuning {
    uint32_t pool_paths;            /* path slots in flight (queue capacity)     */
    uint32_t workgroup;             /* threads per workgroup: 64, 128 or 256     */
    uint32_t groups_per_cu;         /* workgroups per compute unit (grid size)   */
    uint32_t iterations_per_graph;  /* wavefront iterations per hipGraph launch
                                       (even, >= 2)                              */
    uint32_t bvh_width;             /* children per BVH node walked: 4 or 8 (every
                                       resident scene holds both trees)          */
} crg_tuning;
int crg_get_tuning(crg_ctx* ctx, crg_tuning* out);
/* Waits for the requests in flight, then applies the configuration. */
int crg_set_tuning(crg_ctx* ctx, const crg_tuning* tuning);
/* Times short renders of `sample` under candidate configurations, applies the
 * fastest and returns it (with its measured rate in *samples_per_second).  Pool
 * sizes are searched only when the sample has at least half as many work items
 * as the configured pool has slots; otherwise the configured pool is kept. */
int crg_autotune(crg_ctx* ctx, const crg_request* sample, crg_tuning* out, double* samples_per_second);

/* ---- film encode -------------------------------------------------------- */
/*
 * The reference's output writers (reference src/io/image.cpp:201-236, called by
 * src/main.cpp:247-248): io::writePng (exposure in stops, ACES-like filmic
 * curve, sRGB, 8 bits) and io::writeHdr (unexposed linear radiance as Radiance
 * RGBE with run-length coded scanlines).  The film (width*height RGB floats, as
 * crg_render writes it) is checked and encoded on the device.
 *
 * rgb8 (width*height*3 bytes) receives the display bytes writePng stores; hdr
 * (capacity hdr_capacity, at least crg_film_hdr_bound) receives the complete
 * .hdr file writeHdr writes, its length in *hdr_size.  Either may be NULL.
 * Checks and messages follow the writers: dimensions, then pixels (finite,
 * non-negative; for HDR below 2^127), then (PNG) the exposure range.
 */
int crg_film_encode(crg_ctx* ctx, const float* film_rgb, int32_t width, int32_t height, float exposure,
                    uint8_t* rgb8, uint8_t* hdr, uint64_t hdr_capacity, uint64_t* hdr_size);
/* Upper bound on the .hdr file size for a width x height film. */
uint64_t crg_film_hdr_bound(int32_t width, int32_t height);
/* io::writePng / io::writeHdr to a file (reference src/io/image.cpp:201-236).
 * The PNG holds the same pixel bytes as the reference's (its deflate stream
 * is zlib's, not stb's); the .hdr file is byte-identical. */
int crg_write_png(crg_ctx* ctx, const char* path, const float* film_rgb, int32_t width, int32_t height,
                  float exposure);
int crg_write_hdr(crg_ctx* ctx, const char* path, const float* film_rgb, int32_t width, int32_t height);

/* ---- per-function device probe (parity tests) ---------------------------
 * Runs one function of the device path on gfx950 over an array of input
 * records, so each can be compared with the reference's own values.  No
 * counterpart in the reference: test infrastructure for the parity tests of
 * the functions render() is made of.  Records (in -> out), one per element:
 *   CRG_PROBE_CAMERA       float px, py              -> crg_probe_ray  (Camera::generate,
 *                          with `camera`, `width`, `height`; reference src/geometry.cpp:213-234)
 *   CRG_PROBE_ENV_SAMPLE   uint64_t rng state        -> crg_probe_env_sample (Environment::sample)
 *   CRG_PROBE_ENV_PDF      float direction[3]        -> float          (Environment::pdf)
 *   CRG_PROBE_ENV_RADIANCE float direction[3]        -> float rgb[3]   (Environment::radiance)
 *                          (the three with `env_id`; reference src/render.cpp:45-94)
 *   CRG_PROBE_BSDF_SAMPLE  crg_probe_bsdf_in         -> crg_probe_bsdf_sample (sampleBsdf,
 *                          reference src/bsdf.cpp:184-237)
 *   CRG_PROBE_BSDF_EVAL    crg_probe_bsdf_in         -> float value[3], pdf (evaluateBsdf,
 *                          reference src/bsdf.cpp:176-181)
 *   CRG_PROBE_OFFSET_RAY   crg_probe_offset_in       -> crg_probe_ray  (offsetRay,
 *                          reference src/geometry.cpp:186-209)
 *   CRG_PROBE_SURFACE_AT   crg_probe_surface_in      -> crg_surface    (surfaceAt over `scene_id`,
 *                          reference src/geometry.cpp:401-459; tangent/bitangent are not
 *                          computed: the device's BSDF does not read them) */
#define CRG_PROBE_CAMERA 0
#define CRG_PROBE_ENV_SAMPLE 1
#define CRG_PROBE_ENV_PDF 2
#define CRG_PROBE_ENV_RADIANCE 3
#define CRG_PROBE_BSDF_SAMPLE 4
#define CRG_PROBE_BSDF_EVAL 5
#define CRG_PROBE_OFFSET_RAY 6
#define CRG_PROBE_SURFACE_AT 7

/* The reference's Surface (reference src/scene.h:105-115). */
typedef struct crg_surface {
    float position[3], position_low[3], geometric_normal[3], normal[3], tangent[3], bitangent[3];
    float base_color[3], emission[3], attenuation_color[3];
    float metallic, roughness, transmission, ior, thickness, attenuation_distance;
    uint32_t material;
    int32_t front_face, double_sided;
} crg_surface;

typedef struct crg_probe_ray {
    crg_ray ray;     /* origin, direction, tmin, tmax (origin_low and ignore zero) */
    int32_t status;  /* 0, or the failure (camera: invalid; offset: invalid / overflow) */
    int32_t pad;
} crg_probe_ray;

typedef struct crg_probe_env_sample {
    float direction[3], pdf;
    uint64_t rng_end;  /* generator state after the draws */
} crg_probe_env_sample;

typedef struct crg_probe_bsdf_in {
    crg_surface surface;
    float wo[3], wi[3];
    uint64_t rng;      /* generator state (sample) */
} crg_probe_bsdf_in;

/* The reference's BsdfSample (reference src/bsdf.h:19-29); valid = operator bool. */
typedef struct crg_probe_bsdf_sample {
    float direction[3], weight[3];
    float pdf, eta;
    int32_t delta, transmitted, valid;
    uint64_t rng_end;
} crg_probe_bsdf_sample;

typedef struct crg_probe_offset_in {
    float position[3], geometric_normal[3], direction[3], max_distance;
} crg_probe_offset_in;

typedef struct crg_probe_surface_in {
    crg_ray ray;
    crg_hit hit;  /* triangle: finalized index; t, u, v: binary64 */
} crg_probe_surface_in;

int crg_device_probe(crg_ctx* ctx, int32_t what, uint64_t scene_id, uint64_t env_id, const crg_camera* camera,
                     int32_t width, int32_t height, const void* in, uint64_t count, void* out);

/* ---- measurement ------------------------------------------------------- */
/* Per-kernel device time, measured with HIP events recorded on the context's
 * stream around every launch while profiling is enabled.  No counterpart in
 * the reference (it has no device); used by the benchmark's roofline figures. */
typedef struct crg_kernel_time {
    char name[32];
    uint64_t launches;
    double total_ms;
} crg_kernel_time;

int crg_set_profiling(crg_ctx* ctx, int enable);
/* Traversal kernels that count their work per query (the crg_stats work
 * counters) for the requests that follow; off by default, so the timed
 * kernels carry no counters.  Waits for the requests in flight. */
int crg_set_work_counters(crg_ctx* ctx, int enable);
/* Copies up to `max` entries; *count receives the number of kernels tracked. */
int crg_kernel_times(crg_ctx* ctx, crg_kernel_time* out, int max, int* count);

#ifdef __cplusplus
}
#endif
#endif /* CRGPU_H */
