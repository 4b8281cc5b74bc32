// crgpu_render.hpp — cr::gpu::render with exactly the parameter list of the
// reference's cr::render (reference src/render.h:53), on device 0.
//
// Include it where the reference's render.h is included (it includes that
// header itself) and change `cr::render(` to `cr::gpu::render(`: the scene,
// BVH, camera, environment, options and progress arguments and the
// RenderResult are the reference's own.  The BVH argument is not read (the
// library builds its own device tree from the scene).  A constant Environment
// converts directly; an image Environment throws std::invalid_argument (its
// texels are private), so image lighting goes through the overload of
// include/crgpu_cr.hpp that takes a crg_env_desc.
#pragma once
#include "render.h"
#include "crgpu_cr.hpp"

namespace cr::gpu {

// The context every call of the drop-in below uses: device 0, created on first
// use and kept for the life of the process.
inline Device& default_device() {
    static Device dev(0);
    return dev;
}

inline RenderResult render(const Scene& scene, const Bvh& bvh, const Camera& camera, const Environment& environment,
                           const RenderOptions& options, Progress progress = {}) {
    return render<RenderResult>(default_device(), scene, bvh, camera, environment, options, std::move(progress));
}

}  // namespace cr::gpu
