"""MI355X (gfx950) wavefront path tracer: host-side Python mirror of the
reference renderer's render()/Bvh API over the crgpu C ABI."""
