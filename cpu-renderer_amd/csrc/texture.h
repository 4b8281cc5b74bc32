#pragma once
// Texture sampling shared by traversal (alpha-mask acceptance) and shading.
// Scene::sample (reference src/geometry.cpp:238-275): the UV transform and the
// wrap run in emulated binary64 like the reference's doubles; texels are
// filtered in float.  Texels are linear RGBA float4 rows, image after image.

#include "layout.h"

namespace cr {


// wrappedCoordinate (reference src/geometry.cpp:77-84).
CR_HDI sf::f64 wrap_coord(sf::f64 x, int wrap) {
    const sf::f64 zero = sf::zero(0), one = sf::one(), two = sf::two();
    if (wrap == 1) return sf::lt(x, zero) ? zero : (sf::lt(one, x) ? one : x);
    if (wrap == 0) return sf::sub(x, sf::floor_(x));
    sf::f64 r = sf::fmod2(x);
    if (sf::lt(r, zero)) r = sf::add(r, two);
    return sf::le(r, one) ? r : sf::sub(two, r);
}

// wrappedIndex (reference src/geometry.cpp:87-95).
CR_HDI int wrap_index(int64_t i, int size, int wrap) {
    const int64_t n = size;
    if (wrap == 1) return int(i < 0 ? 0 : (i > n - 1 ? n - 1 : i));
    const int64_t period = wrap == 2 ? 2 * n : n;
    i %= period;
    if (i < 0) i += period;
    if (wrap == 2 && i >= n) i = period - 1 - i;
    return int(i);
}

CR_HDI float4 mix4(float4 a, float4 b, float t) {
    return make_float4(a.x * (1 - t) + b.x * t, a.y * (1 - t) + b.y * t, a.z * (1 - t) + b.z * t,
                       a.w * (1 - t) + b.w * t);
}

struct UvPair {
    float u0[2], u1[2];
};

// interpolateUv for both sets (reference src/geometry.cpp:68-74).
CR_HDI UvPair interp_uv(const DevScene& sc, uint32_t prim, sf::f64 hu, sf::f64 hv) {
    const sf::f64 w = sf::sub(sf::sub(sf::one(), hu), hv);
    const float4 a = sc.uvs[3 * prim], b = sc.uvs[3 * prim + 1], c = sc.uvs[3 * prim + 2];
    const float av[4] = {a.x, a.y, a.z, a.w}, bv[4] = {b.x, b.y, b.z, b.w}, cv[4] = {c.x, c.y, c.z, c.w};
    float out[4];
    for (int k = 0; k < 4; ++k)
        out[k] = sf::to_float(sf::add(sf::add(sf::mul(w, sf::from_float(av[k])), sf::mul(hu, sf::from_float(bv[k]))),
                                      sf::mul(hv, sf::from_float(cv[k]))));
    return UvPair{{out[0], out[1]}, {out[2], out[3]}};
}

// Returns (1,1,1,1) for an unbound slot; `bad` is set for non-finite UVs.
CR_OUTLINE float4 sample_tex(const DevScene& sc, const DevSlot& slot, const UvPair& uv, bool srgb, bool& bad) {
    if (slot.texture < 0) return make_float4(1, 1, 1, 1);
    const DevTexture tex = sc.textures[slot.texture];
    const DevImage im = sc.images[tex.image];
    const float* t = slot.texcoord == 0 ? uv.u0 : uv.u1;
    if (!finite1(t[0]) || !finite1(t[1])) {
        bad = true;
        return make_float4(1, 1, 1, 1);
    }
    const sf::f64 c{slot.cos_bits}, sn{slot.sin_bits};
    const sf::f64 x = sf::mul(sf::from_float(t[0]), sf::from_float(slot.scale[0]));
    const sf::f64 y = sf::mul(sf::from_float(t[1]), sf::from_float(slot.scale[1]));
    const sf::f64 u =
        wrap_coord(sf::add(sf::sub(sf::mul(c, x), sf::mul(sn, y)), sf::from_float(slot.offset[0])), tex.wrap_s);
    const sf::f64 v =
        wrap_coord(sf::add(sf::add(sf::mul(sn, x), sf::mul(c, y)), sf::from_float(slot.offset[1])), tex.wrap_t);
    // Colour slots read the image's sRGB-decoded copy, decoded per texel on
    // the host with its libm (the reference decodes each fetched texel).
    const uint64_t base = srgb ? im.srgb_offset : im.offset;
    auto texel = [&](int64_t ix, int64_t iy) {
        const int tx = wrap_index(ix, im.width, tex.wrap_s), ty = wrap_index(iy, im.height, tex.wrap_t);
        return sc.texels[base + uint64_t(ty) * uint64_t(im.width) + uint64_t(tx)];
    };
    const sf::f64 w = sf::from_i32(im.width), h = sf::from_i32(im.height);
    if (!tex.linear)
        return texel(sf::to_i64(sf::floor_(sf::mul(u, w))), sf::to_i64(sf::floor_(sf::mul(v, h))));
    const sf::f64 half{0x3fe0000000000000ull};
    const sf::f64 px = sf::sub(sf::mul(u, w), half), py = sf::sub(sf::mul(v, h), half);
    const sf::f64 fpx = sf::floor_(px), fpy = sf::floor_(py);
    const int64_t ix = sf::to_i64(fpx), iy = sf::to_i64(fpy);
    const float fx = sf::to_float(sf::sub(px, fpx)), fy = sf::to_float(sf::sub(py, fpy));
    return mix4(mix4(texel(ix, iy), texel(ix + 1, iy), fx), mix4(texel(ix, iy + 1), texel(ix + 1, iy + 1), fx), fy);
}

// True when a bound slot would read non-finite texture coordinates (the
// reference throws from Scene::sample then).
CR_HDI bool slot_uv_bad(const DevSlot& slot, const UvPair& uv) {
    if (slot.texture < 0) return false;
    const float* t = slot.texcoord == 0 ? uv.u0 : uv.u1;
    return !finite1(t[0]) || !finite1(t[1]);
}

// The MASK part of accepts() (reference src/geometry.cpp:279-292): base-colour
// alpha times the sampled texel alpha (no sRGB decode) against the cutoff, at
// the binary64 barycentrics of the hit.
CR_HDI bool mask_accepts(const DevScene& sc, uint32_t prim, uint32_t material, sf::f64 hu, sf::f64 hv, bool& bad) {
    const DevMaterial& m = sc.materials[material];
    const float4 c = sample_tex(sc, m.base_tex, interp_uv(sc, prim, hu, hv), false, bad);
    return m.base[3] * c.w >= m.alpha_cutoff;
}

}  // namespace cr
