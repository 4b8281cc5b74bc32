// Instantiates include/crgpu_cr.hpp with the reference renderer's own types (CPU test).
#include "render.h"
#include "crgpu_cr.hpp"
#include <cstdio>
int main() {
  cr::Scene s; cr::Material m; s.materials.push_back(m);
  cr::Triangle t; t.v[1].position = {1,0,0}; t.v[2].position = {0,1,0}; s.triangles.push_back(t);
  cr::Bvh bvh; bvh.build(s); cr::Camera cam; cr::RenderOptions o; o.width = 4; o.height = 4; o.samples = 1;
  crg_env_desc env{}; env.constant[0] = env.constant[1] = env.constant[2] = 0.5f; env.intensity = 1;
  try { cr::gpu::Device dev(0);
    auto r = cr::gpu::render<cr::RenderResult>(dev, s, bvh, cam, env, o, cr::Progress{});
    std::printf("rendered %zu px\n", r.pixels.size());
  } catch (const std::exception& e) { std::printf("device: %s\n", e.what()); }
  return 0;
}
