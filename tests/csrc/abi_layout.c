/* Prints sizeof/offsetof of the crgpu.h descriptors as JSON (CPU ABI test). */
#include <stddef.h>
#include <stdio.h>
#include "../../include/crgpu.h"
#define S(T) printf("\"%s\": %zu,\n", #T, sizeof(T))
int main(void) {
    printf("{\n");
    S(crg_texture_slot); S(crg_material); S(crg_light); S(crg_image); S(crg_texture); S(crg_scene_desc);
    S(crg_camera); S(crg_env_desc); S(crg_options); S(crg_stats); S(crg_ray); S(crg_hit); S(crg_scene_info);
    S(crg_kernel_time); S(crg_request); S(crg_tuning); S(crg_engine_info);
    printf("\"crg_stats.exact_fallbacks\": %zu,\n", offsetof(crg_stats, exact_fallbacks));
    printf("\"crg_stats.latency_seconds\": %zu,\n", offsetof(crg_stats, latency_seconds));
    printf("\"crg_request.film_rgb\": %zu,\n", offsetof(crg_request, film_rgb));
    printf("\"crg_options.seed\": %zu\n}\n", offsetof(crg_options, seed));
    return 0;
}
