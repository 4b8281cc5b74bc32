// host_scene.h — host-side scene preparation for the device (C++20).
//
// Everything here runs once per scene/environment/camera on the CPU, in
// native binary64 where the reference uses double, so the tables the device
// consumes are bit-identical to the values the reference computes per call.
#pragma once
#include <cstdint>
#include <stdexcept>
#include <string>
#include <vector>

#include "../../include/crgpu.h"
#include "layout.h"

namespace cr {

struct HVec3 {
    float x, y, z;
};

struct HTriangle {
    HVec3 p[3];
    HVec3 n[3];
    float tan[3][4];  // finalized tangent.xyz and handedness per vertex
    float uv[3][4];   // uv0.xy, uv1.xy per vertex
    uint32_t material;
    HVec3 ng;   // geometric normal (finalize)
    HVec3 ung;  // unit(ng) in binary64 (offsetRay)
    float area;
};

struct HImage {
    int width = 0, height = 0;
    std::vector<float4> texels;
};

struct HostScene {
    std::vector<HTriangle> tris;
    std::vector<crg_material> materials;
    std::vector<crg_light> lights;
    std::vector<HImage> images;
    std::vector<crg_texture> textures;
    std::vector<std::string> warnings;
};

// Validate and repair imported triangles (reference src/geometry.cpp:296-397).
HostScene finalize_scene(const crg_scene_desc& desc);

struct WideBvh {
    std::vector<WNode> nodes;   // 4-wide float tree (host)
    std::vector<QNode4> q4;     // device: the same tree with quantized child boxes
    std::vector<QNode8> q8;     // device: 8-wide collapse of it over the same leaves
    std::vector<uint32_t> order;  // prim i -> finalized triangle index (a split triangle has several prims)
    uint64_t binary_nodes = 0;    // equivalent binary node count (2*leaves-1)
};

// Binned SAH (16 bins, leaves <= 4, depth <= 48) emitted as 4-wide nodes,
// then quantized (q4) and collapsed to 8-wide nodes (q8).
WideBvh build_wide_bvh(const HostScene& s);
// q4 and q8 of b.nodes.
void quantize_trees(WideBvh& b);
// unit() of the reference in binary64, rounded to float (reference src/geometry.cpp:31-37).
void unit_vector(const float* v, float* out);
// One quantized 4-wide node over n <= 4 child boxes (lo/hi: 3 floats per child).
QNode4 quantize_boxes4(const float* lo, const float* hi, const uint32_t* refs, int n);

// Environment-independent part of the light mixture of one scene.
struct SceneLights {
    double radius2 = 1;
    std::vector<double> punctual_weight;  // per punctual light
    std::vector<uint32_t> emissive;       // finalized triangle indices with weight > 0
    std::vector<double> emissive_weight;
    std::vector<DevPunctual> punctual;
};

struct LightTable {
    std::vector<DevLightEntry> entries;
    float env_probability = 0;
};

struct HostEnv {
    bool is_image = false;
    float constant[3] = {0, 0, 0};
    int width = 0, height = 0;
    float rotation = 0, intensity = 1;
    double integral = 0, weight_sum = 0;
    std::vector<float> texels;  // rgba
    std::vector<float> pdf;
    std::vector<uint32_t> threshold;
    std::vector<uint32_t> guide;
    std::vector<float> rowcos;
    std::vector<double> cdf;  // the reference's cumulative weights (host-side checks)
};

// Environment construction (reference src/render.cpp:12-42, :67-94).
HostEnv build_env(const crg_env_desc& d);

// Lights mixture (reference src/render.cpp:131-160): weights per scene, then the
// normalized table for a given environment integral.  tri_to_prim maps a
// finalized triangle index to its device primitive.
SceneLights scene_lights(const HostScene& s);
LightTable build_light_table(const SceneLights& sl, const std::vector<uint32_t>& tri_to_prim, double env_integral);
// Per-prim index of its entry in any table built from `sl` (kNoTriangle if not emissive).
// Light rank of every leaf-order primitive (kNoTriangle: not emissive).
std::vector<uint32_t> light_ranks(const SceneLights& sl, const std::vector<uint32_t>& prim_to_tri, size_t triangles);

// Camera basis and raster constants (reference src/geometry.cpp:213-227).
DevCamera prepare_camera(const crg_camera& c, int width, int height);

// RenderOptions::validate (reference src/render.cpp:97-104).
void validate_options(const crg_options& o);

// "Pixel (x,y), sample s: <cause>" for a device error record, worded like the
// reference's rethrown exceptions (reference src/render.cpp:437-438, :310-323).
std::string format_path_error(const ErrorRec& e, int width, int spp);

// Material table with precomputed absorption (reference src/bsdf.cpp:29-34)
// and texture slots with their rotation's binary64 cos/sin.
std::vector<DevMaterial> device_materials(const HostScene& s);

// Everything the device reads for one scene, in BVH leaf order.
struct SceneArrays {
    WideBvh bvh;
    std::vector<uint32_t> tri_to_prim;
    std::vector<PrimRec> prims;
    std::vector<float4> prim_ng;
    std::vector<PrimAux> aux;
    std::vector<DevMaterial> materials;
    std::vector<float4> uvs;       // empty unless a material samples a texture
    std::vector<float4> tangents;  // empty unless a material has a normal map
    std::vector<float4> texels;
    std::vector<DevImage> images;
    std::vector<DevTexture> textures;
};
SceneArrays scene_arrays(const HostScene& hs);

// (cos, sin)(2*pi*k/2^24) for every 24-bit draw k, with this host's cosf/sinf
// on the reference's float expression (reference src/bsdf.cpp, src/render.cpp
// sampling sites); out holds kTurnEntries entries.
void fill_turn_table(TurnCS* out);

}  // namespace cr
