// shade.h — per-hit transport pieces of the wavefront integrator (device).
//
// Each function is a restatement of one reference routine, evaluated in the
// same float expression order so that per-sample results agree bit for bit
// wherever the math library agrees; binary64 parts use the sf64 emulation:
//   camera_ray      Camera::generate        reference src/geometry.cpp:213-234
//   offset_ray      offsetRay               reference src/geometry.cpp:186-209
//   surface_at      surfaceAt (no textures) reference src/geometry.cpp:401-459
//   Bsdf*           Context/evalLocal/sampleBsdf/evaluateBsdf  reference src/bsdf.cpp:6-237
//   env_*           Environment             reference src/render.cpp:12-94
//   sample_light    Lights::sample          reference src/render.cpp:172-228
//   triangle_pdf    Lights::trianglePdf     reference src/render.cpp:165-170
#pragma once
#include "layout.h"
#include "texture.h"

namespace cr {

// ---------------------------------------------------------------- camera
// Camera::generate with every binary64 step emulated.
CR_OUTLINE bool camera_ray_chain(const DevCamera& c, float px, float py, V3& o, V3& d, float& tmin, float& tmax) {
    const sf::f64 one = sf::one(), two = sf::two();
    const sf::f64 x = sf::mul(
        sf::mul(sf::sub(sf::div(sf::mul(two, sf::from_float(px)), sf::from_i32(c.width)), one), c.ratio),
        c.half_height);
    const sf::f64 y =
        sf::mul(sf::sub(one, sf::div(sf::mul(two, sf::from_float(py)), sf::from_i32(c.height))), c.half_height);
    const V3 f{c.f[0], c.f[1], c.f[2]}, r{c.r[0], c.r[1], c.r[2]}, u{c.u[0], c.u[1], c.u[2]};
    const D3 dir = dadd(dadd(widen(f), dscale(widen(r), x)), dscale(widen(u), y));
    d = dunit_ff(dir, V3{0, 0, 0});
    const sf::f64 proj = ddot(widen(d), widen(f));
    const sf::f64 near_t = sf::div(sf::from_float(c.near_clip), proj);
    if (!sf::gt(proj, sf::zero(0)) || !sf::is_finite(near_t) || sf::gt(near_t, sf::from_float(kFltMax))) return false;
    o = V3{c.position[0], c.position[1], c.position[2]};
    tmin = sf::to_float(near_t);
    tmax = sf::to_float(sf::div(sf::from_float(c.far_clip), proj));
    return true;
}

// hi + lo rounded to float when the reference's binary64 value (within
// `slack` relative of hi + lo) rounds to the same float: at least
// 2^-38 + slack (relative) from a binary32 rounding boundary, in the normal range.
CR_HDI bool ff_round(float hi, float lo, float slack, float& out) {
    const float f = hi + lo, af = fabsf(f);
    const float d = fabsf((hi - f) + lo);
    const float half_ulp = 0.5f * fminf(nextafterf(af, INFINITY) - af, af - nextafterf(af, 0.0f));
    if (!(af >= 0x1p-100f && af <= 0x1p+126f) || !(half_ulp - d >= (0x1p-38f + slack) * af)) return false;
    out = f;
    return true;
}

// num / den for a float-pair den, as a float pair to about 2^-46 relative.
CR_HDI FF ff_div(float num, FF den) {
    const float q0 = num / den.hi;
    const float r = (fmaf(-q0, den.hi, num)) - q0 * den.lo;
    return ff_fast_two_sum(q0, r / den.hi);
}

// Camera::generate in float-float (the binary64 chain below covers the rest):
// x = (2 px / width - 1) ratio halfHeight and y = (1 - 2 py / height) halfHeight
// as pairs with absolute error bounds that cover the reference's binary64
// roundings (the subtraction of 1 can cancel), the direction through unit_ff
// with those bounds, the projection on f from exact float products, and the
// clip parameters rounded through ff_round.  False where the chain decides,
// including every ray the reference rejects.
CR_HDI bool camera_ray_ff(const DevCamera& c, float px, float py, V3& d, float& tmin, float& tmax) {
    FF ratio, hh;
    if (!ff_of(c.ratio, ratio) || !ff_of(c.half_height, hh) || !(fabsf(px) < 0x1p+60f) || !(fabsf(py) < 0x1p+60f) ||
        c.width <= 0 || c.width > (1 << 24) || c.height <= 0 || c.height > (1 << 24))
        return false;
    const float rh = fabsf(ratio.hi) * fabsf(hh.hi);
    const FF qx = ff_div(2.0f * px, FF{float(c.width), 0.0f}), qy = ff_div(2.0f * py, FF{float(c.height), 0.0f});
    const FF xm = ff_add(qx, FF{-1.0f, 0.0f}), ym = ff_add(FF{1.0f, 0.0f}, FF{-qy.hi, -qy.lo});
    const FF x = ff_mul(ff_mul(xm, ratio), hh), y = ff_mul(ym, hh);
    const float ex = 0x1p-44f * (fabsf(qx.hi) + fabsf(xm.hi)) * rh + 0x1p-120f;
    const float ey = 0x1p-44f * (fabsf(qy.hi) + fabsf(ym.hi)) * fabsf(hh.hi) + 0x1p-120f;
    FF v[3];
    float err[3];
    for (int k = 0; k < 3; ++k) {
        const float fk = c.f[k], rk = c.r[k], uk = c.u[k];
        const FF t1 = ff_mul(FF{rk, 0.0f}, x), t2 = ff_mul(FF{uk, 0.0f}, y);
        v[k] = ff_add(ff_add(FF{fk, 0.0f}, t1), t2);
        err[k] = fabsf(rk) * ex + fabsf(uk) * ey + 0x1p-44f * ((fabsf(fk) + fabsf(t1.hi)) + fabsf(t2.hi)) + 0x1p-120f;
    }
    float dd[3];
    if (!unit_ff(v, err, dd)) return false;
    FF proj{0.0f, 0.0f};
    float mag = 0.0f;
    for (int k = 0; k < 3; ++k) {
        const float p = dd[k] * c.f[k];
        proj = ff_add(proj, FF{p, fmaf(dd[k], c.f[k], -p)});
        mag += fabsf(p);
    }
    const float slack = 0x1p-44f * mag / proj.hi;
    if (!(proj.hi > 0.0f) || !(slack < 0x1p-20f)) return false;
    float t0, t1;
    if (c.near_clip == 0.0f) t0 = c.near_clip;
    else {
        const FF nt = ff_div(c.near_clip, proj);
        if (!ff_round(nt.hi, nt.lo, slack, t0)) return false;
    }
    if (c.far_clip == INFINITY) t1 = INFINITY;
    else {
        const FF ft = ff_div(c.far_clip, proj);
        if (!ff_round(ft.hi, ft.lo, slack, t1)) return false;
    }
    d = V3{dd[0], dd[1], dd[2]};
    tmin = t0;
    tmax = t1;
    return true;
}

CR_OUTLINE bool camera_ray(const DevCamera& c, float px, float py, V3& o, V3& d, float& tmin, float& tmax) {
    if (camera_ray_ff(c, px, py, d, tmin, tmax)) {
        o = V3{c.position[0], c.position[1], c.position[2]};
        return true;
    }
    return camera_ray_chain(c, px, py, o, d, tmin, tmax);
}

// ---------------------------------------------------------------- offsetRay
// One coordinate of offsetRay's shifted point, float(p + distance * n) with
// distance = 2^-21 * scale: the binary64 sum of the reference (the product is
// exact in binary64) rounded to float.  In float-float: scale * n as an exact
// pair scaled by 2^-21, its sum with p to about 2^-47 relative, one rounding;
// within 2^-38 (relative) of a binary32 rounding boundary, or where the pair's
// low part could leave the normal range, the binary64 emulation decides (the
// distance is formed there, off the common path).  A zero n adds a zero of
// n's sign, as the binary64 sum does.
CR_HDI float offset_coord(float p, float scale, float n) {
    if (n == 0.0f) return p + n;
    const float ph = scale * n, pl = fmaf(scale, n, -ph);
    const float aph = fabsf(ph);
    if (aph >= 0x1p-50f && aph <= 0x1p+100f && fabsf(p) <= 0x1p+100f) {
        const FF sum = ff_two_sum(p, ph * 0x1p-21f);
        const float lo = sum.lo + pl * 0x1p-21f;
        const float f = sum.hi + lo, af = fabsf(f);
        const float d = fabsf((sum.hi - f) + lo);
        const float half_ulp = 0.5f * fminf(nextafterf(af, INFINITY) - af, af - nextafterf(af, 0.0f));
        if (af >= 0x1p-100f && half_ulp - d >= 0x1p-38f * af) return f;
    }
    const sf::f64 distance = sf::mul(sf::from_float(4 * kFltEps), sf::from_float(scale));
    return sf::to_float(sf::add(sf::from_float(p), sf::mul(distance, sf::from_float(n))));
}

// `ung` is unit(geometricNormal) (binary64 normalization done once per triangle
// on the host); `ng` is the stored geometric normal used for the validity test.
CR_OUTLINE uint32_t offset_ray(V3 p, V3 ng, V3 ung, V3 direction, float max_distance, V3& origin, V3& dir,
                           float& tmax) {
    if (!finite(p) || !finite(ng) || !finite(direction) || !nonzero(ng) || !nonzero(direction) ||
        max_distance != max_distance || max_distance < 0)
        return kErrInvalidSecondary;
    dir = dunit_ff(widen(direction), V3{0, 0, 0});
    V3 n = ung;
    if (dot_sign_exact(n, dir) < 0) n = -n;
    float scale = fabsf(p.x);
    if (scale < fabsf(p.y)) scale = fabsf(p.y);
    if (scale < fabsf(p.z)) scale = fabsf(p.z);
    for (int a = 0; a < 3; ++a) {
        const float na = at(n, a);
        float v = offset_coord(at(p, a), scale, na);
        if (na != 0) v = na > 0 ? next_up(v) : next_down(v);
        set(origin, a, v);
    }
    if (!finite(origin)) return kErrOffsetOverflow;
    float limit = max_distance;
    if (finite1(limit)) {
        const sf::f64 advance = ddot(dsub(widen(origin), widen(p)), widen(dir));
        limit = sf::to_float(sf::max_(sf::zero(0), sf::sub(sf::from_float(limit), advance)));
        limit = next_toward_zero(limit);
    }
    tmax = limit;
    return 0;
}

// ---------------------------------------------------------------- surfaceAt
struct Surf {
    V3 position, position_low, ng, normal;
    V3 base, emission, att_color;
    float metallic, roughness, transmission, ior, thickness, att_distance;
    uint32_t material;
    bool front, double_sided;
    bool texture_error;  // non-finite texture coordinates (reported like the reference's exception)
};

CR_HDI void material_into(const DevMaterial& m, Surf& s) {
    s.base = V3{m.base[0], m.base[1], m.base[2]};
    s.metallic = clampf(m.metallic, 0.0f, 1.0f);
    s.roughness = clampf(m.roughness, 0.0f, 1.0f);
    s.emission = V3{m.emission[0], m.emission[1], m.emission[2]};
    s.transmission = clampf(m.transmission, 0.0f, 1.0f);
    s.ior = m.ior;
    s.thickness = m.thickness;
    s.att_distance = m.att_distance;
    s.att_color = V3{m.att_color[0], m.att_color[1], m.att_color[2]};
    s.double_sided = m.double_sided != 0;
    s.texture_error = false;
}

// orthogonalized() (reference src/geometry.cpp:40-47) in emulated binary64.
CR_HDI V3 orthogonalize(const D3& v, V3 n, V3 fallback) {
    const D3 dn = widen(n);
    const sf::f64 nn = ddot(dn, dn);
    if (!sf::gt(nn, sf::zero(0))) return fallback;
    const D3 proj = dsub(v, dscale(dn, sf::div(ddot(v, dn), nn)));
    const sf::f64 tol{0x3ea0000000000000ull};  // 4 * FLT_EPSILON = 2^-21
    if (sf::le(ddot(proj, proj), sf::mul(sf::mul(ddot(v, v), tol), tol))) return fallback;
    return dunit_ff(proj, fallback);
}

// Material maps and normal mapping of surfaceAt (reference src/geometry.cpp:420-457).
// The tangent frame after mapping is not consumed by the BSDF (it shades in
// Frame(normal)), so only the mapped normal is produced.
CR_OUTLINE void apply_maps(const DevScene& sc, const DevMaterial& m, uint32_t prim, sf::f64 w, sf::f64 hu,
                           sf::f64 hv, Surf& s) {
    bool bad = false;
    const UvPair uv = interp_uv(sc, prim, hu, hv);
    const float4 base = sample_tex(sc, m.base_tex, uv, true, bad);
    s.base = V3{m.base[0], m.base[1], m.base[2]} * V3{base.x, base.y, base.z};
    const float4 mr = sample_tex(sc, m.mr_tex, uv, false, bad);
    s.metallic = clampf(m.metallic * mr.z, 0.0f, 1.0f);
    s.roughness = clampf(m.roughness * mr.y, 0.0f, 1.0f);
    const float4 em = sample_tex(sc, m.emissive_tex, uv, true, bad);
    s.emission = V3{m.emission[0], m.emission[1], m.emission[2]} * V3{em.x, em.y, em.z};
    s.transmission = clampf(m.transmission * sample_tex(sc, m.transmission_tex, uv, false, bad).x, 0.0f, 1.0f);
    if (m.normal_tex.texture >= 0) {
        const float4 t0 = sc.tangents[3 * prim], t1 = sc.tangents[3 * prim + 1], t2 = sc.tangents[3 * prim + 2];
        const D3 ti = dadd(dadd(dscale(widen(V3{t0.x, t0.y, t0.z}), w), dscale(widen(V3{t1.x, t1.y, t1.z}), hu)),
                           dscale(widen(V3{t2.x, t2.y, t2.z}), hv));
        V3 tangent = orthogonalize(ti, s.normal, make_frame(s.normal).x);
        const sf::f64 hsum = sf::add(sf::add(sf::mul(w, sf::from_float(t0.w)), sf::mul(hu, sf::from_float(t1.w))),
                                     sf::mul(hv, sf::from_float(t2.w)));
        float hand = sf::lt(hsum, sf::zero(0)) ? -1.0f : 1.0f;
        V3 bitangent = cross(s.normal, tangent) * hand;
        const DevSlot& sl = m.normal_tex;
        if (sl.scale[0] != 0 && sl.scale[1] != 0) {
            const sf::f64 c{sl.cos_bits}, sn{sl.sin_bits};
            const D3 tt = dsub(dscale(widen(tangent), sf::div(c, sf::from_float(sl.scale[0]))),
                               dscale(widen(bitangent), sf::div(sn, sf::from_float(sl.scale[1]))));
            tangent = orthogonalize(tt, s.normal, tangent);
            if ((sl.scale[0] < 0) != (sl.scale[1] < 0)) hand = -hand;
            bitangent = cross(s.normal, tangent) * hand;
        }
        const float4 tx = sample_tex(sc, sl, uv, false, bad);
        const sf::f64 one = sf::one(), two = sf::two(), ns = sf::from_float(m.normal_scale);
        const sf::f64 kx = sf::mul(sf::sub(sf::mul(two, sf::from_float(tx.x)), one), ns);
        const sf::f64 ky = sf::mul(sf::sub(sf::mul(two, sf::from_float(tx.y)), one), ns);
        const sf::f64 kz = sf::sub(sf::mul(two, sf::from_float(tx.z)), one);
        const D3 mapped = dadd(dadd(dscale(widen(tangent), kx), dscale(widen(bitangent), ky)), dscale(widen(s.normal), kz));
        s.normal = dunit_ff(mapped, s.normal);
        if (dot(s.normal, s.ng) < 0) s.normal = -s.normal;
    }
    s.texture_error = bad;
}

// `want_low`: also the binary64 position's low part (precise origins need it).
// Without it the float position alone usually follows from a float-float sum,
// and the binary64 chain runs only where that cannot decide its rounding.
CR_OUTLINE Surf surface_at(const DevScene& sc, uint32_t prim, V3 ray_dir, sf::f64 hu, sf::f64 hv, bool want_low) {
    const PrimRec p = sc.prims[prim];
    const PrimAux& ax = sc.aux[prim];
    Surf s;
    const sf::f64 w = sf::sub(sf::sub(sf::one(), hu), hv);
    const V3 va{p.ax, p.ay, p.az}, vb{p.bx, p.by, p.bz}, vc{p.cx, p.cy, p.cz};
    FF fw, fu, fv;
    const bool have_ff = ff_of(w, fw) && ff_of(hu, fu) && ff_of(hv, fv);
    s.position_low = V3{0, 0, 0};
    if (want_low || !have_ff || !interp_point_ff(va, vb, vc, fw, fu, fv, s.position)) {
        const D3 P = dadd(dadd(dscale(widen(va), w), dscale(widen(vb), hu)), dscale(widen(vc), hv));
        s.position = narrow(P);
        s.position_low = V3{sf::to_float(sf::sub(P.x, sf::from_float(s.position.x))),
                            sf::to_float(sf::sub(P.y, sf::from_float(s.position.y))),
                            sf::to_float(sf::sub(P.z, sf::from_float(s.position.z)))};
    }
    s.ng = V3{ax.ng[0], ax.ng[1], ax.ng[2]};
    s.front = dot_sign_exact(ray_dir, s.ng) < 0;
    const V3 n0{ax.n0[0], ax.n0[1], ax.n0[2]}, n1{ax.n1[0], ax.n1[1], ax.n1[2]}, n2{ax.n2[0], ax.n2[1], ax.n2[2]};
    if (!have_ff || !interp_unit_ff(n0, n1, n2, fw, fu, fv, s.normal))
        s.normal = dunit_ff(dadd(dadd(dscale(widen(n0), w), dscale(widen(n1), hu)), dscale(widen(n2), hv)), s.ng);
    if (dot(s.normal, s.ng) < 0) s.normal = -s.normal;
    s.material = p.material;
    const DevMaterial& m = sc.materials[p.material];
    material_into(m, s);
    if (m.maps) apply_maps(sc, m, prim, w, hu, hv, s);
    return s;
}

// surface_at from float-pair barycentrics with error bounds (hit_uv_ff), for
// materials without maps and queries without a split origin: the position and
// the shading normal where their float rounding is decided; false otherwise
// (the caller then runs the binary64 chain and surface_at).
CR_OUTLINE bool surface_at_ff(const DevScene& sc, uint32_t prim, V3 ray_dir, const HitUvFF& h, Surf& s) {
    const PrimRec p = sc.prims[prim];
    const PrimAux& ax = sc.aux[prim];
    const V3 va{p.ax, p.ay, p.az}, vb{p.bx, p.by, p.bz}, vc{p.cx, p.cy, p.cz};
    if (!interp_point_ff(va, vb, vc, h.w, h.u, h.v, s.position, h.ew, h.eu, h.ev)) return false;
    const V3 n0{ax.n0[0], ax.n0[1], ax.n0[2]}, n1{ax.n1[0], ax.n1[1], ax.n1[2]}, n2{ax.n2[0], ax.n2[1], ax.n2[2]};
    if (!interp_unit_ff(n0, n1, n2, h.w, h.u, h.v, s.normal, h.ew, h.eu, h.ev)) return false;
    s.position_low = V3{0, 0, 0};
    s.ng = V3{ax.ng[0], ax.ng[1], ax.ng[2]};
    s.front = dot_sign_exact(ray_dir, s.ng) < 0;
    if (dot(s.normal, s.ng) < 0) s.normal = -s.normal;
    s.material = p.material;
    material_into(sc.materials[p.material], s);
    return true;
}

// ---------------------------------------------------------------- sampling angles
// cos and sin of phi = 2*pi*u for the 24-bit draw k (u = k/2^24) come from a
// table the host fills with its own cosf/sinf on the reference's float
// expression (2 * pi * u): the device's cosf/sinf round differently on about
// one input in five, and a direction off by one ulp can change which triangle
// a ray hits.
#if defined(__HIPCC__)
static __device__ const TurnCS* g_turn_table;
#else
inline const TurnCS* g_turn_table = nullptr;
#endif
CR_HDI void turn_cos_sin(uint32_t k, float& c, float& s) {
    const TurnCS t = g_turn_table[k];
    c = t.c;
    s = t.s;
}

// ---------------------------------------------------------------- BSDF
CR_HDI float fresnel_dielectric(float cosine, float eta_i, float eta_t) {
    if (eta_i == eta_t) return 0;
    cosine = clampf(fabsf(cosine), 0.0f, 1.0f);
    const float eta = eta_i / eta_t;
    const float sin_t2 = eta * eta * fmax_(0.0f, 1 - cosine * cosine);
    if (sin_t2 >= 1) return 1;
    const float cos_t = sqrtf(1 - sin_t2);
    const float ds = eta_i * cosine + eta_t * cos_t, dp = eta_t * cosine + eta_i * cos_t;
    if (ds == 0 || dp == 0) return 1;
    const float rs = (eta_i * cosine - eta_t * cos_t) / ds, rp = (eta_t * cosine - eta_i * cos_t) / dp;
    return 0.5f * (rs * rs + rp * rp);
}

CR_HDI bool refract_direction(V3 wo, V3 n, float ratio, V3& wi) {
    const float cos_i = dot(wo, n);
    const float sin_t2 = ratio * ratio * fmax_(0.0f, 1 - cos_i * cos_i);
    if (sin_t2 >= 1) return false;
    wi = normalize(-wo * ratio + n * (ratio * cos_i - sqrtf(1 - sin_t2)));
    return true;
}

// x^5 for the Schlick term: powf(x, 5) correctly rounded, from an exact
// binary64 product chain (x^5 of a float needs at most 120 bits; the chain
// below rounds once per step at 53 bits, far inside the float rounding margin).
CR_HDI float pow5(float x) {
    const sf::f64 d = sf::from_float(x);
    const sf::f64 d2 = sf::mul(d, d);
    return sf::to_float(sf::mul(sf::mul(d2, d2), d));
}

// The same in float-float: x^2 exact (fma), x^4 and x^5 to about 2^-45
// relative, one rounding; within 2^-36 (relative) of a binary32 rounding
// boundary, and outside the normal range, the binary64 chain decides.
CR_HDI float pow5_ff(float x) {
    const float ax = fabsf(x);
    if (!(ax >= 0x1p-12f && ax <= 0x1p+24f)) return pow5(x);  // low parts would leave the normal range; NaN
    const float h2 = x * x, l2 = fmaf(x, x, -h2);
    const float h4 = h2 * h2, l4 = fmaf(h2, h2, -h4) + 2.0f * h2 * l2;
    const float h5 = h4 * x, l5 = fmaf(h4, x, -h5) + l4 * x;
    const float f = h5 + l5, af = fabsf(f);
    const float d = fabsf((h5 - f) + l5);
    const float half_ulp = 0.5f * fminf(nextafterf(af, INFINITY) - af, af - nextafterf(af, 0.0f));
    if (half_ulp - d < 0x1p-36f * af) return pow5(x);
    return f;
}

CR_HDI float ggx_d(float alpha, V3 h) {
    if (h.z <= 0) return 0;
    const float a2 = alpha * alpha;
    const float d = (h.x * h.x + h.y * h.y) + a2 * h.z * h.z;
    return a2 / (kPi * d * d);
}

CR_HDI float smith_g1(float alpha, V3 v) {
    const float cos2 = v.z * v.z;
    if (cos2 == 0) return 0;
    return 2 / (1 + sqrtf(1 + alpha * alpha * (v.x * v.x + v.y * v.y) / cos2));
}

CR_HDI V3 visible_normal(V3 wo, float alpha, float u, uint32_t kv) {
    const V3 vh = normalize(V3{alpha * wo.x, alpha * wo.y, wo.z});
    const float lensq = vh.x * vh.x + vh.y * vh.y;
    const V3 t1 = lensq > 0 ? V3{-vh.y, vh.x, 0} / sqrtf(lensq) : V3{1, 0, 0};
    const V3 t2 = cross(vh, t1);
    const float r = sqrtf(u);
    float cp, sp;
    turn_cos_sin(kv, cp, sp);
    const float x = r * cp;
    float y = r * sp;
    const float s = 0.5f * (1 + vh.z);
    y = (1 - s) * sqrtf(fmax_(0.0f, 1 - x * x)) + s * y;
    const V3 nh = t1 * x + t2 * y + vh * sqrtf(fmax_(0.0f, 1 - x * x - y * y));
    return normalize(V3{alpha * nh.x, alpha * nh.y, fmax_(0.0f, nh.z)});
}

struct Bsdf {
    const Surf* s;
    Frame frame;
    V3 wo;
    float eta_i, eta_t, eta, alpha, diffuse_p, delta_t_p;
    bool thin, smooth, matched;

    CR_HDI float fresnel(float cosine) const {
        const float f = fresnel_dielectric(cosine, eta_i, eta_t);
        return thin ? 2 * f / (1 + f) : f;
    }
    CR_HDI V3 reflection(float f, float cosine) const {
        const float schlick = pow5_ff(1 - clampf(fabsf(cosine), 0.0f, 1.0f));
        return v3((1 - s->metallic) * f) + (s->base + (v3(1) - s->base) * schlick) * s->metallic;
    }
    CR_HDI V3 transmission(float f) const { return s->base * ((1 - s->metallic) * s->transmission * (1 - f)); }
    CR_HDI float reflection_probability(float f, float cos_h) const {
        if (matched) return 1;
        const float r = luminance(reflection(f, cos_h)), t = luminance(transmission(f));
        return r + t > 0 ? r / (r + t) : 1;
    }
    CR_HDI float micro_probability() const { return fmax_(0.0f, 1 - diffuse_p - delta_t_p); }
};

CR_HDI void bsdf_init(Bsdf& c, const Surf& s, V3 outgoing) {
    c.s = &s;
    c.frame = make_frame(dot(outgoing, s.normal) > 0 ? s.normal : -s.normal);
    c.wo = to_local(c.frame, outgoing);
    const bool exit = s.thickness > 0 && !s.front;
    c.eta_i = exit ? s.ior : 1;
    c.eta_t = exit ? 1 : s.ior;
    c.eta = c.eta_t / c.eta_i;
    c.alpha = fmax_(0.001f, s.roughness * s.roughness);
    c.diffuse_p = 0;
    c.delta_t_p = 0;
    c.thin = s.thickness <= 0 && s.transmission > 0;
    c.smooth = s.roughness <= 0.001f;
    c.matched = false;
    if (c.thin) {
        c.eta_i = 1;
        c.eta_t = s.ior;
        c.eta = s.ior;
    }
    c.matched = fabsf(c.eta_i - c.eta_t) <= sqrtf(kFltEps) * fmax_(c.eta_i, c.eta_t);
    if (c.matched) c.eta_i = c.eta_t = c.eta = 1;
    const float f = c.fresnel(c.wo.z);
    const float diffuse = (1 - s.metallic) * (1 - s.transmission) * luminance(s.base) * (1 - f);
    const float transmit = (1 - s.metallic) * s.transmission * (1 - f) * luminance(s.base);
    const float total = diffuse + luminance(c.reflection(f, c.wo.z)) + transmit;
    if (total > 0) {
        c.diffuse_p = diffuse / total;
        if (c.matched) c.delta_t_p = transmit / total;
    }
}

struct BsdfEval {
    V3 value;
    float pdf;
};

CR_HDI BsdfEval eval_local(const Bsdf& c, V3 wi) {
    BsdfEval r{V3{0, 0, 0}, 0};
    const float no = c.wo.z, ni = fabsf(wi.z);
    if (no <= 1e-7f || ni <= 1e-7f) return r;
    const bool refl = wi.z > 0;
    if (!refl && (c.matched || c.thin)) return r;
    const Surf& s = *c.s;
    if (refl) {
        const float diffuse = (1 - s.metallic) * (1 - s.transmission) * (1 - c.fresnel(no)) * (1 - c.fresnel(ni));
        r.value = s.base * (diffuse / kPi);
        r.pdf = c.diffuse_p * ni / kPi;
    }
    if (c.smooth) return r;
    V3 h = normalize(refl ? c.wo + wi : c.wo + wi * c.eta);
    if (h.z < 0) h = -h;
    const float oh = dot(c.wo, h), ih = dot(wi, h);
    if (oh <= 0 || (refl ? ih <= 0 : ih >= 0)) return r;
    const float f = c.fresnel(oh), pr = c.reflection_probability(f, oh);
    const float d = ggx_d(c.alpha, h), g1 = smith_g1(c.alpha, c.wo), g = g1 * smith_g1(c.alpha, wi);
    const float pdf_h = d * g1 * oh / no;
    if (refl) {
        r.value = r.value + c.reflection(f, oh) * (d * g / (4 * no * ni));
        r.pdf += c.micro_probability() * pr * pdf_h / (4 * oh);
    } else {
        const float denom = oh + c.eta * ih, denom2 = denom * denom;
        if (denom2 < 1e-20f) return BsdfEval{V3{0, 0, 0}, 0};
        r.value = c.transmission(f) * (d * g * fabsf(oh * ih) / (no * ni * denom2));
        r.pdf = c.micro_probability() * (1 - pr) * pdf_h * c.eta * c.eta * fabsf(ih) / denom2;
    }
    return r;
}

CR_HDI bool correct_side(const Surf& s, V3 wo, V3 wi, bool transmitted) {
    const float d = dot(wo, s.ng) * dot(wi, s.ng);
    return transmitted ? d < 0 : d > 0;
}

CR_OUTLINE BsdfEval evaluate_bsdf(const Surf& s, V3 wo, V3 wi) {
    Bsdf c;
    bsdf_init(c, s, wo);
    const V3 local = to_local(c.frame, wi);
    if (!correct_side(s, wo, wi, local.z < 0)) return BsdfEval{V3{0, 0, 0}, 0};
    return eval_local(c, local);
}

struct BsdfSample {
    V3 direction, weight;
    float pdf, eta;
    bool delta, transmitted;
};

CR_HDI BsdfSample empty_sample() { return BsdfSample{V3{0, 0, 0}, V3{0, 0, 0}, 0, 1, false, false}; }
CR_HDI bool sample_ok(const BsdfSample& b) {
    return max_component(b.weight) > 0 && finite(b.weight) && finite(b.direction);
}

CR_OUTLINE BsdfSample sample_bsdf(const Surf& s, V3 wo, Rng& rng) {
    Bsdf c;
    bsdf_init(c, s, wo);
    BsdfSample out = empty_sample();
    if (c.wo.z <= 1e-7f) return out;
    V3 wi;
    const float lobe = rng.uniform();
    if (lobe < c.diffuse_p) {
        const float u = rng.uniform();
        float cp, sp;
        turn_cos_sin(rng.bits24(), cp, sp);
        wi = V3{sqrtf(u) * cp, sqrtf(u) * sp, sqrtf(1 - u)};
    } else if (lobe < c.diffuse_p + c.delta_t_p) {
        out.direction = -wo;
        out.delta = true;
        out.transmitted = true;
        out.pdf = c.delta_t_p;
        out.weight = s.base * ((1 - s.metallic) * s.transmission / out.pdf);
        return correct_side(s, wo, out.direction, true) ? out : empty_sample();
    } else {
        const float u = rng.uniform();
        const uint32_t kv = rng.bits24();
        const V3 h = c.smooth ? V3{0, 0, 1} : visible_normal(c.wo, c.alpha, u, kv);
        const float oh = dot(c.wo, h), f = c.fresnel(oh), pr = c.reflection_probability(f, oh);
        const bool refl = rng.uniform() < pr;
        if (refl) wi = reflect(-c.wo, h);
        else if (c.thin) wi = -c.wo;
        else if (!refract_direction(c.wo, h, 1 / c.eta, wi)) return out;
        if ((refl && wi.z <= 0) || (!refl && wi.z >= 0)) return out;
        out.transmitted = !refl;
        out.eta = out.transmitted && !c.thin ? c.eta : 1;
        if (c.smooth || (!refl && c.thin)) {
            out.delta = true;
            out.direction = normalize(to_world(c.frame, wi));
            if (!correct_side(s, wo, out.direction, out.transmitted)) return empty_sample();
            const float prob = c.micro_probability() * (refl ? pr : 1 - pr);
            if (prob <= 0) return empty_sample();
            out.pdf = prob;
            out.weight = (refl ? c.reflection(f, oh) : c.transmission(f) / (out.eta * out.eta)) / prob;
            return out;
        }
    }
    out.direction = normalize(to_world(c.frame, wi));
    wi = to_local(c.frame, out.direction);
    if ((wi.z < 0) != out.transmitted || !correct_side(s, wo, out.direction, out.transmitted)) return empty_sample();
    const BsdfEval e = eval_local(c, wi);
    if (e.pdf <= 0 || !finite1(e.pdf)) return empty_sample();
    out.pdf = e.pdf;
    out.weight = e.value * (fabsf(wi.z) / e.pdf);
    return out;
}

// e^x correctly rounded to binary32, as sf::expf_cr, in float-float arithmetic:
// x = n*ln2 + r with ln2 split into two 16-bit heads (exact products for
// |n| < 2^8) and a float tail, r carried as an unevaluated float pair; the
// Taylor series of e^r to r^10 in float-float Horner form (truncation and
// rounding below 2^-40 relative for |r| <= ln2/2).  The pair is rounded once;
// when its value lies within 2^-38 (relative) of a binary32 rounding boundary,
// or the result would leave the normal range, the binary64 emulation decides.
CR_HDI float expf_ff(float x) {
    if (!(x >= -86.0f && x <= 88.0f)) return sf::expf_cr(x);  // NaN, underflow, overflow edges
    const float n = rintf(x * 1.44269502163f);
    const float a = x - n * 0x1.62e4p-1f;  // exact (Sterbenz)
    FF r = ff_two_sum(a, -(n * 0x1.7f7ep-20f));
    r = ff_two_sum(r.hi, r.lo - n * -0x1.c610cap-37f);
    constexpr float kHi[11] = {0x1p+0f,        0x1p+0f,        0x1p-1f,         0x1.555556p-3f,
                               0x1.555556p-5f, 0x1.111112p-7f, 0x1.6c16c2p-10f, 0x1.a01a02p-13f,
                               0x1.a01a02p-16f, 0x1.71de3ap-19f, 0x1.27e4fcp-22f};
    constexpr float kLo[11] = {0.0f,           0.0f,           0.0f,            -0x1.555556p-28f,
                               -0x1.555556p-30f, -0x1.dddddep-32f, -0x1.27d27ep-35f, -0x1.7f97fap-39f,
                               -0x1.7f97fap-42f, 0x1.55b1ccp-45f, -0x1.10ec14p-47f};
    FF p{kHi[10], kLo[10]};
    for (int k = 9; k >= 0; --k) p = ff_add(ff_mul(p, r), FF{kHi[k], kLo[k]});
    const float f = p.hi + p.lo;
    const float d = fabsf((p.hi - f) + p.lo);              // distance of the pair from f
    const float half_ulp = 0.5f * fminf(nextafterf(f, 4.0f) - f, f - nextafterf(f, 0.0f));
    if (half_ulp - d < 0x1p-38f * f) return sf::expf_cr(x);
    return ldexpf(f, int(n));
}

// powerHeuristic (reference src/math.h:124-128) without the binary64 division:
// a^2 and b^2 exact as float pairs, their sum and the quotient to about 2^-45
// relative in float-float, one rounding.  The reference's binary64 quotient is
// within 2^-51 of the exact one, so both round alike unless the value lies
// within 2^-38 (relative) of a binary32 rounding boundary; there, and for
// arguments whose squares would leave the normal range, the binary64 chain
// (power_heuristic) decides.
CR_HDI float power_heuristic_ff(float a, float b) {
    if (a <= 0) return 0;
    const float ab = fabsf(b);
    if (!(a >= 0x1p-40f && a <= 0x1p+40f && (ab == 0.0f || (ab >= 0x1p-40f && ab <= 0x1p+40f))))
        return power_heuristic(a, b);
    const float ah = a * a, al = fmaf(a, a, -ah);
    const float bh = b * b, bl = fmaf(b, b, -bh);
    const FF sum = ff_add(FF{ah, al}, FF{bh, bl});
    const float q0 = ah / sum.hi;
    const float p = q0 * sum.hi, pe = fmaf(q0, sum.hi, -p);
    const float r = (((ah - p) - pe) + al) - q0 * sum.lo;
    const float q1 = r / sum.hi;
    const float f = q0 + q1, af = fabsf(f);
    const float d = fabsf((q0 - f) + q1);
    const float half_ulp = 0.5f * fminf(nextafterf(af, INFINITY) - af, af - nextafterf(af, 0.0f));
    if (!(af >= 0x1p-100f) || !(half_ulp - d >= 0x1p-38f * af)) return power_heuristic(a, b);
    return f;
}

// Beer-Lambert transmittance over `distance` (reference src/bsdf.cpp:37-40).
CR_HDI V3 beer_lambert(V3 sigma, float distance) {
    if (distance == kInf) return V3{sigma.x > 0 ? 0.0f : 1.0f, sigma.y > 0 ? 0.0f : 1.0f, sigma.z > 0 ? 0.0f : 1.0f};
    const V3 a = sigma * (-fmax_(0.0f, distance));
    // expf correctly rounded: agrees with the host libm's expf on all but about
    // 1 in 10^5 arguments (the device's expf differs on about 1 in 16).
    const float ex = expf_ff(a.x);
    if (a.y == a.x && a.z == a.x) return V3{ex, ex, ex};
    return V3{ex, expf_ff(a.y), expf_ff(a.z)};
}

// ---------------------------------------------------------------- environment
CR_HDI void env_uv(const DevEnv& e, V3 d, float& u, float& v) {
    u = (atan2f(d.z, d.x) - e.rotation) / (2 * kPi) + 0.5f;
    u -= floorf(u);
    v = acosf(clampf(d.y, -1.0f, 1.0f)) / kPi;
}

CR_HDI V3 env_texel(const DevEnv& e, int ix, int iy) {
    ix = (ix % e.width + e.width) % e.width;
    iy = iy < 0 ? 0 : (iy > e.height - 1 ? e.height - 1 : iy);
    const float4 t = e.texels[size_t(iy) * size_t(e.width) + size_t(ix)];
    return V3{t.x, t.y, t.z};
}

CR_OUTLINE V3 env_radiance(const DevEnv& e, V3 d) {
    if (!e.is_image) return V3{e.constant[0], e.constant[1], e.constant[2]};
    float u, v;
    env_uv(e, d, u, v);
    const float px = u * float(e.width) - 0.5f, py = v * float(e.height) - 0.5f;
    const int x = int(floorf(px)), y = int(floorf(py));
    const float tx = px - float(x), ty = py - float(y);
    return mix(mix(env_texel(e, x, y), env_texel(e, x + 1, y), tx),
               mix(env_texel(e, x, y + 1), env_texel(e, x + 1, y + 1), tx), ty) *
           e.intensity;
}

CR_OUTLINE float env_pdf(const DevEnv& e, V3 d) {
    if (!e.is_image) return 1 / (4 * kPi);
    float u, v;
    env_uv(e, d, u, v);
    int x = int(u * float(e.width));
    if (x > e.width - 1) x = e.width - 1;
    int y = int(v * float(e.height));
    if (y > e.height - 1) y = e.height - 1;
    return e.pdf[size_t(y) * size_t(e.width) + size_t(x)];
}

// Texel index for a 24-bit draw k: the count of thresholds T_i <= k, i.e. the
// reference's upper_bound over its binary64 CDF (clamped to the last texel).
CR_HDI uint32_t env_invert(const DevEnv& e, uint32_t k) {
    const uint32_t n = uint32_t(e.width) * uint32_t(e.height);
    uint32_t lo = e.guide[k >> kGuideShift], hi = e.guide[(k >> kGuideShift) + 1];
    while (lo < hi) {  // first index in [lo, hi) with T > k
        const uint32_t mid = (lo + hi) >> 1;
        if (e.threshold[mid] <= k) lo = mid + 1;
        else hi = mid;
    }
    return lo < n ? lo : n - 1;
}

CR_OUTLINE V3 env_sample(const DevEnv& e, Rng& rng, float& density) {
    if (!e.is_image) {
        const float y = 1 - 2 * rng.uniform();
        float cp, sp;
        turn_cos_sin(rng.bits24(), cp, sp);
        const float r = sqrtf(fmax_(0.0f, 1 - y * y));
        density = 1 / (4 * kPi);
        return V3{r * cp, y, r * sp};
    }
    const uint32_t index = env_invert(e, rng.bits24());
    const int x = int(index % uint32_t(e.width)), y = int(index / uint32_t(e.width));
    const float u = (float(x) + rng.uniform()) / float(e.width);
    const float c0 = e.rowcos[y], c1 = e.rowcos[y + 1];
    const float cos_t = c0 + (c1 - c0) * rng.uniform(), sin_t = sqrtf(fmax_(0.0f, 1 - cos_t * cos_t));
    const float phi = (u - 0.5f) * 2 * kPi + e.rotation;
    density = e.pdf[index];
    return V3{sin_t * cosf(phi), cos_t, sin_t * sinf(phi)};
}

// ---------------------------------------------------------------- lights
struct LightSample {
    V3 direction, radiance, target, target_normal, target_unormal;
    float distance, pdf;
    bool delta;
    bool tex_error;  // the emitter's textures read non-finite coordinates
};

CR_HDI LightSample no_light() {
    LightSample l;
    l.direction = l.radiance = l.target = l.target_normal = l.target_unormal = V3{0, 0, 0};
    l.distance = kInf;
    l.pdf = 0;
    l.delta = false;
    l.tex_error = false;
    return l;
}

CR_HDI float triangle_pdf(const DevScene& sc, const DevLights& lt, uint32_t prim, V3 from, V3 point) {
    const PrimAux& ax = sc.aux[prim];
    const V3 ng{ax.ng[0], ax.ng[1], ax.ng[2]};
    const V3 dir = point - from;
    const uint32_t rank = sc.light_rank[prim];
    const float prob = rank == kNoTriangle ? 0.0f : lt.entries[rank].probability;
    const float len2 = length_sq(dir), cosine = fabsf(dot(ng, normalize(dir)));
    return cosine > 1e-8f ? prob * len2 / (ax.area * cosine) : 0;
}

CR_HDI float environment_pdf(const DevLights& lt, const DevEnv& e, V3 d) { return lt.env_probability * env_pdf(e, d); }

CR_OUTLINE LightSample sample_light(const DevScene& sc, const DevLights& lt, const DevEnv& env, V3 from, Rng& rng) {
    LightSample out = no_light();
    if (lt.count == 0) return out;
    const uint32_t k = rng.bits24();
    uint32_t lo = 0, hi = lt.count;  // first entry whose threshold >= k
    while (lo < hi) {
        const uint32_t mid = (lo + hi) >> 1;
        if (lt.entries[mid].threshold < k) lo = mid + 1;
        else hi = mid;
    }
    const DevLightEntry le = lt.entries[lo < lt.count ? lo : lt.count - 1];
    if (le.type == 2) {
        out.direction = env_sample(env, rng, out.pdf);
        out.pdf *= le.probability;
        out.radiance = env_radiance(env, out.direction);
        return out;
    }
    if (le.type == 0) {
        const DevPunctual& l = sc.punctual[le.index];
        out.delta = true;
        out.pdf = le.probability;
        float attenuation = 1;
        if (l.type == 1) {
            out.direction = -V3{l.direction[0], l.direction[1], l.direction[2]};
        } else {
            const V3 lp{l.position[0], l.position[1], l.position[2]};
            const V3 delta = lp - from;
            const float r2 = length_sq(delta);
            if (r2 < 1e-12f) return no_light();
            out.distance = sqrtf(r2);
            out.direction = delta / out.distance;
            out.target = lp;
            attenuation = 1 / r2;
            if (finite1(l.range)) {
                const float q = out.distance / l.range;
                attenuation *= fmax_(0.0f, 1 - q * q * q * q);
            }
            if (l.type == 2) {
                const V3 ld{l.direction[0], l.direction[1], l.direction[2]};
                const float cos_a = dot(ld, -out.direction);
                const float falloff =
                    clampf((cos_a - l.cos_outer) / fmax_(1e-7f, l.cos_inner - l.cos_outer), 0.0f, 1.0f);
                attenuation *= falloff * falloff;
            }
        }
        out.radiance = V3{l.color[0], l.color[1], l.color[2]} * (l.intensity * attenuation);
        return out;
    }
    // Emissive triangle: uniform area sample by the square-root warp.
    const uint32_t prim = le.index;
    const PrimRec p = sc.prims[prim];
    const PrimAux& ax = sc.aux[prim];
    const float r = sqrtf(rng.uniform()), u1 = 1 - r, u2 = r * rng.uniform(), u0 = 1 - u1 - u2;
    out.target = V3{p.ax, p.ay, p.az} * u0 + V3{p.bx, p.by, p.bz} * u1 + V3{p.cx, p.cy, p.cz} * u2;
    out.target_normal = V3{ax.ng[0], ax.ng[1], ax.ng[2]};
    out.target_unormal = V3{ax.ung[0], ax.ung[1], ax.ung[2]};
    const V3 delta = out.target - from;
    out.distance = length(delta);
    if (out.distance < 1e-7f) return no_light();
    out.direction = delta / out.distance;
    if (p.flags & kPrimNeverHit) return no_light();
    const int side = dot_sign_exact(out.target_normal, out.direction);
    if ((p.flags & kPrimCull) && side >= 0) return no_light();
    const DevMaterial& m = sc.materials[p.material];
    // accepts() and surfaceAt() at the sampled barycentrics (u1, u2).
    V3 emission{m.emission[0], m.emission[1], m.emission[2]};
    if (m.maps) {
        const sf::f64 hu = sf::from_float(u1), hv = sf::from_float(u2);
        const UvPair uv = interp_uv(sc, prim, hu, hv);
        bool bad = slot_uv_bad(m.base_tex, uv);
        if (!bad && (p.flags & kPrimMaskTexture) && !mask_accepts(sc, prim, p.material, hu, hv, bad))
            return no_light();
        bad = bad || slot_uv_bad(m.mr_tex, uv) || slot_uv_bad(m.emissive_tex, uv) ||
              slot_uv_bad(m.transmission_tex, uv) || slot_uv_bad(m.normal_tex, uv);
        if (bad) {
            LightSample e = no_light();
            e.tex_error = true;
            return e;
        }
        const float4 em = sample_tex(sc, m.emissive_tex, uv, true, bad);
        emission = emission * V3{em.x, em.y, em.z};
    }
    const bool front = side < 0;
    if (!front && !m.double_sided) return no_light();
    out.radiance = emission;
    out.pdf = triangle_pdf(sc, lt, prim, from, out.target);
    return out;
}

}  // namespace cr
