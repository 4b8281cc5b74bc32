/*
 * probe.h — TEST INFRASTRUCTURE ONLY.
 *
 * Function-level probe interface implemented twice:
 *   - oracle/ref_harness.cpp  (prefix ref_) links the reference's own sources,
 *     compiled from where they lie, into oracle/_ref/libcrref.so;
 *   - oracle/cr_oracle.cpp    (prefix orc_) is the CPU restatement of the
 *     reference algorithm, built into oracle/liboracle.so.
 * Only tests/, __graft_entry__.smoke() and bench.py's cpu_baseline leg load
 * either library, and only as the checker — never as the product path.
 */
#ifndef CR_PROBE_H
#define CR_PROBE_H
#include "../include/crgpu.h"

#ifdef __cplusplus
extern "C" {
#endif

/* Mirror of the reference Surface (reference src/scene.h:105-115). */
typedef struct probe_surface {
    float position[3], position_low[3], geometric_normal[3], normal[3], tangent[3], bitangent[3];
    float base_color[3], emission[3], attenuation_color[3];
    float metallic, roughness, transmission, ior, thickness, attenuation_distance;
    uint32_t material;
    int32_t front_face, double_sided;
} probe_surface;

/* Mirror of the reference BsdfSample (reference src/bsdf.h:19-29); valid = operator bool. */
typedef struct probe_bsdf_sample {
    float direction[3], weight[3];
    float pdf, eta;
    int32_t delta, transmitted, valid;
} probe_bsdf_sample;

#define PROBE_DECLARE(P)                                                                              \
    const char* P##_last_error(void);                                                                 \
    uint64_t P##_sample_seed(uint64_t seed, uint64_t pixel, uint64_t sample);                         \
    void P##_rng_draw(uint64_t* state, int count, float* uniforms);                                   \
    int P##_camera_generate(const crg_camera* camera, float px, float py, int width, int height,      \
                            crg_ray* out);                                                            \
    void* P##_scene_create(const crg_scene_desc* desc);                                               \
    void P##_scene_destroy(void* scene);                                                              \
    uint64_t P##_scene_triangles(void* scene);                                                        \
    uint64_t P##_scene_nodes(void* scene);                                                            \
    int P##_trace(void* scene, const crg_ray* rays, uint64_t count, crg_hit* hits, int any);           \
    int P##_offset_ray(const float* position, const float* normal, const float* direction,            \
                       float max_distance, crg_ray* out);                                             \
    int P##_surface_at(void* scene, const crg_ray* ray, const crg_hit* hit, probe_surface* out);       \
    void P##_bsdf_sample(const probe_surface* s, const float* wo, uint64_t* rng,                      \
                         probe_bsdf_sample* out);                                                     \
    void P##_bsdf_eval(const probe_surface* s, const float* wo, const float* wi, float* value_pdf);   \
    void* P##_env_create(const crg_env_desc* desc);                                                   \
    void P##_env_destroy(void* env);                                                                  \
    void P##_env_sample(void* env, uint64_t* rng, float* direction, float* pdf);                      \
    float P##_env_pdf(void* env, const float* direction);                                             \
    void P##_env_radiance(void* env, const float* direction, float* rgb);                             \
    int P##_render(void* scene, void* env, const crg_camera* camera, const crg_options* options,      \
                   float* film, crg_stats* stats);                                                     \
    int P##_film_png8(const float* film, int width, int height, float exposure, uint8_t* rgb8);       \
    int P##_film_hdr(const float* film, int width, int height, uint8_t* out, uint64_t capacity,       \
                     uint64_t* size);

PROBE_DECLARE(ref)
PROBE_DECLARE(orc)

/* Oracle only: render() over a window (x0, y0, w, h) of the frame with global
 * pixel indices (the full-frame sample streams); film is w*h RGB, samples
 * (NULL or spp*w*h RGB, [sample][window pixel]) the radiance of every sample. */
int orc_render_window(void* scene, void* env, const crg_camera* camera, const crg_options* options, int x0, int y0,
                      int w, int h, float* film, float* samples, crg_stats* stats);

/* Reference only: io::writePng and io::writeHdr to the given paths (timing). */
int ref_film_write(const float* film, int width, int height, float exposure, const char* png_path,
                   const char* hdr_path);

/* Reference only: io::loadGltf (reference src/io/gltf.cpp:743-879) flattened into a
 * crgpu scene descriptor by include/crgpu_cr.hpp's describe(), so real assets can be
 * written out as fixtures; render_loaded renders the loaded cr::Scene itself. */
void* ref_gltf_load(const char* path);
void ref_gltf_destroy(void* gltf);
const crg_scene_desc* ref_gltf_desc(void* gltf);
int ref_gltf_camera_count(void* gltf);
int ref_gltf_camera(void* gltf, int index, crg_camera* out);
int ref_gltf_render_loaded(void* gltf, void* env, const crg_camera* camera, const crg_options* options,
                           float* film, crg_stats* stats);

#ifdef __cplusplus
}
#endif
#endif
