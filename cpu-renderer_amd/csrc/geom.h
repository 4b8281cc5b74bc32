// geom.h — ray/triangle intersection and 4-wide BVH traversal (device).
//
// Closest hit is the lexicographic minimum of (t, triangle id) over accepted
// hits with t in [tmin, tmax], t and the hit predicate being the reference's
// binary64 watertight test (reference src/geometry.cpp:147-182, ordering at
// :609, acceptance rules at :279-292).  The result does not depend on tree
// shape, so the device uses its own 4-wide tree and visiting order.
//
// Each candidate is first tested in FP32 together with a rigorous forward
// error bound that covers both the FP32 and the reference's binary64 rounding.
// When the bound decides every edge sign and the t-interval checks, the FP32
// decision equals the reference's; otherwise (rays grazing an edge, near-ties
// in t) the triangle is re-evaluated with the exact binary64 emulation in the
// reference's expression order.  Box tests are FP32 slabs widened so they can
// only over-include.
#pragma once
#include "layout.h"
#include "texture.h"

namespace cr {

constexpr float kU = 5.9604644775390625e-8f;  // 2^-24, unit roundoff of binary32

struct RayQuery {
    V3 o, d, olow;
    float tmin, tmax;
    uint32_t ignore;
    int kx, ky, kz;
    float sx, sy, sz;
    float inv[3];
    float slack;    // absolute widening of box intervals for split origins
    float ovec;     // 2.02*|olow|_inf: bound on the split-origin part of FP32 vertex offsets
    bool precise;   // origin_low != 0
    bool valid;
    bool have_exact, have_esz;
    sf::f64 esx, esy, esz;  // binary64 shear constants, computed on first use
};

CR_HDI bool valid_ray(V3 o, V3 olow, V3 d, float tmin, float tmax) {
    return finite(o) && finite(olow) && finite(d) && nonzero(d) && tmin == tmin && tmax == tmax && tmin <= tmax;
}

CR_HDI void setup_query(RayQuery& q, V3 o, V3 d, float tmin, float tmax, V3 olow, uint32_t ignore) {
    q.o = o;
    q.d = d;
    q.olow = olow;
    q.tmin = tmin;
    q.tmax = tmax;
    q.ignore = ignore;
    q.valid = valid_ray(o, olow, d, tmin, tmax);
    q.precise = nonzero(olow);
    q.have_exact = false;
    q.have_esz = false;
    int kz = 0;
    if (fabsf(d.y) > fabsf(at(d, kz))) kz = 1;
    if (fabsf(d.z) > fabsf(at(d, kz))) kz = 2;
    int kx = (kz + 1) % 3, ky = (kx + 1) % 3;
    if (at(d, kz) < 0) {
        const int t = kx;
        kx = ky;
        ky = t;
    }
    q.kx = kx;
    q.ky = ky;
    q.kz = kz;
    const float dz = at(d, kz);
    q.sx = at(d, kx) / dz;
    q.sy = at(d, ky) / dz;
    q.sz = 1.0f / dz;
    float slack = 0;
    for (int a = 0; a < 3; ++a) {
        const float c = at(d, a);
        float iv = 1.0f / c;
        // A reciprocal that overflows (subnormal component) would turn finite
        // slab distances into infinities; NaN makes that axis unconstrained.
        if (c != 0 && !finite1(iv)) iv = __builtin_nanf("");
        q.inv[a] = iv;
        if (q.precise && c != 0 && finite1(iv)) slack = fmaxf(slack, fabsf(at(olow, a) * iv) * 2.0f);
    }
    q.slack = slack;
    q.ovec = fmaxf(fabsf(olow.x), fmaxf(fabsf(olow.y), fabsf(olow.z))) * 2.02f;
}

CR_HDI void ensure_exact_shear(RayQuery& q) {
    if (q.have_exact) return;
    const sf::f64 dz = sf::from_float(at(q.d, q.kz));
    q.esx = sf::div(sf::from_float(at(q.d, q.kx)), dz);
    q.esy = sf::div(sf::from_float(at(q.d, q.ky)), dz);
    q.have_exact = true;
}

CR_HDI void ensure_exact_esz(RayQuery& q) {
    if (q.have_esz) return;
    q.esz = sf::div(sf::one(), sf::from_float(at(q.d, q.kz)));
    q.have_esz = true;
}

struct ExactHit {
    bool hit;
    sf::f64 t, u, v;
};

// Origin component o + olow in binary64 when olow is a zero: o itself, with
// the sign of a zero sum (-0 only when both are -0).
CR_HDI float origin_of_zero_low(float o, float olow) {
    return o == 0.0f ? sf::bitsf(sf::fbits(o) & sf::fbits(olow) & 0x80000000u) : o;
}

// Vertex minus ray origin in binary64, components permuted to (kx, ky, kz).
// Without a low origin part every component is a difference of two floats.
CR_HDI void rel_vertex(const RayQuery& q, V3 v, const D3& origin, sf::f64& x, sf::f64& y, sf::f64& z) {
    if (!q.precise) {
        x = sf::sub_floats(at(v, q.kx), origin_of_zero_low(at(q.o, q.kx), at(q.olow, q.kx)));
        y = sf::sub_floats(at(v, q.ky), origin_of_zero_low(at(q.o, q.ky), at(q.olow, q.ky)));
        z = sf::sub_floats(at(v, q.kz), origin_of_zero_low(at(q.o, q.kz), at(q.olow, q.kz)));
        return;
    }
    const D3 a = dsub(widen(v), origin);
    x = dat(a, q.kx);
    y = dat(a, q.ky);
    z = dat(a, q.kz);
}

// The reference's sheared edge functions (reference src/geometry.cpp:154-176):
// the z offsets of the three vertices, w0..w2 and det = w0 + w1 + w2, each in
// binary64 in the reference's expression order.
struct EdgeFunctions {
    sf::f64 az, bz, cz, w0, w1, w2, det;
};

CR_HDI EdgeFunctions edge_functions(RayQuery& q, V3 va, V3 vb, V3 vc) {
    ensure_exact_shear(q);
    D3 origin{sf::zero(0), sf::zero(0), sf::zero(0)};
    if (q.precise) origin = dadd(widen(q.o), widen(q.olow));
    EdgeFunctions e;
    sf::f64 ax, ay, bx, by, cx, cy;
    rel_vertex(q, va, origin, ax, ay, e.az);
    rel_vertex(q, vb, origin, bx, by, e.bz);
    rel_vertex(q, vc, origin, cx, cy, e.cz);
    ax = sf::sub(ax, sf::mul(q.esx, e.az));
    ay = sf::sub(ay, sf::mul(q.esy, e.az));
    bx = sf::sub(bx, sf::mul(q.esx, e.bz));
    by = sf::sub(by, sf::mul(q.esy, e.bz));
    cx = sf::sub(cx, sf::mul(q.esx, e.cz));
    cy = sf::sub(cy, sf::mul(q.esy, e.cz));
    e.w0 = sf::sub(sf::mul(cx, by), sf::mul(cy, bx));
    e.w1 = sf::sub(sf::mul(ax, cy), sf::mul(ay, cx));
    e.w2 = sf::sub(sf::mul(bx, ay), sf::mul(by, ax));
    e.det = sf::add(sf::add(e.w0, e.w1), e.w2);
    return e;
}

// t = (w0*(sz*az) + w1*(sz*bz) + w2*(sz*cz)) / det in binary64.
CR_HDI sf::f64 edge_t(RayQuery& q, const EdgeFunctions& e) {
    ensure_exact_esz(q);
    const sf::f64 num = sf::add(sf::add(sf::mul(e.w0, sf::mul(q.esz, e.az)), sf::mul(e.w1, sf::mul(q.esz, e.bz))),
                                sf::mul(e.w2, sf::mul(q.esz, e.cz)));
    return sf::div(num, e.det);
}

// The reference's intersectTriangle in emulated binary64, same expression order.
CR_OUTLINE ExactHit exact_triangle(RayQuery& q, V3 va, V3 vb, V3 vc, sf::f64 limit_in) {
    ExactHit r;
    r.hit = false;
    if (!q.valid) return r;
    const sf::f64 tmin = sf::from_float(q.tmin);
    const sf::f64 limit = sf::min_(sf::from_float(q.tmax), limit_in);
    if (sf::lt(limit, tmin)) return r;
    const EdgeFunctions e = edge_functions(q, va, vb, vc);
    const sf::f64 z = sf::zero(0);
    const bool neg = sf::lt(e.w0, z) || sf::lt(e.w1, z) || sf::lt(e.w2, z);
    const bool pos = sf::gt(e.w0, z) || sf::gt(e.w1, z) || sf::gt(e.w2, z);
    if (neg && pos) return r;
    if (sf::is_zero(e.det) || !sf::is_finite(e.det)) return r;
    const sf::f64 t = edge_t(q, e);
    if (!sf::is_finite(t) || sf::gt(sf::fabs_(t), sf::from_float(kFltMax)) || sf::lt(t, tmin) || sf::gt(t, limit))
        return r;
    r.hit = true;
    r.t = t;
    r.u = sf::div(e.w1, e.det);
    r.v = sf::div(e.w2, e.det);
    return r;
}

// Barycentrics (and t when `want_t`) of a hit the traversal has already
// decided: the same binary64 values exact_triangle returns for it, without
// its acceptance checks and, unless asked, without t's terms.
CR_OUTLINE ExactHit exact_hit_uv(RayQuery& q, V3 va, V3 vb, V3 vc, bool want_t) {
    const EdgeFunctions e = edge_functions(q, va, vb, vc);
    ExactHit r;
    r.hit = true;
    r.t = want_t ? edge_t(q, e) : sf::zero(0);
    r.u = sf::div(e.w1, e.det);
    r.v = sf::div(e.w2, e.det);
    return r;
}

// Barycentrics of a hit the traversal decided, as float pairs with bounds on
// their distance from the reference's binary64 values (the chain of
// exact_hit_uv): u = w1/det, v = w2/det and w = (1 - u) - v formed from the
// pairs u and v, so that w errs by minus their errors plus at most ew.
struct HitUvFF {
    FF u, v, w;
    float eu, ev, ew;
    // For hit_t_ff: the sheared z offsets' pairs, the edge functions, det and bounds.
    FF az[3], W[3], det;
    float eW[3], edet;
};

// The edge functions in float-float with a running error bound.  Per step the
// pair's arithmetic errs by at most 2^-46 and the reference's binary64 step by
// 2^-53 of the step's operand magnitudes; kStep = 2^-45 covers both with room
// (the vertex offsets are exact pairs; the reference's may round once).  An
// input error propagates through a product as (|x| + ex) ey + |y| ex and
// through a quotient as (e_num + |q| e_den) / (|den| - e_den).  False for
// split-origin queries and where the bound cannot separate det from zero;
// the binary64 chain decides there.
CR_HDI bool hit_uv_ff(const RayQuery& q, V3 va, V3 vb, V3 vc, HitUvFF& h) {
    constexpr float kStep = 0x1p-45f, kFloor = 0x1p-110f;
    if (q.precise || !q.valid) return false;
    const int kx = q.kx, ky = q.ky, kz = q.kz;
    const float dz = at(q.d, kz);
    const FF sx = ff_div2(FF{at(q.d, kx), 0.0f}, FF{dz, 0.0f}), sy = ff_div2(FF{at(q.d, ky), 0.0f}, FF{dz, 0.0f});
    const float ox = at(q.o, kx), oy = at(q.o, ky), oz = at(q.o, kz);
    const V3 vs[3] = {va, vb, vc};
    FF X[3], Y[3];
    float eX[3], eY[3];
#pragma unroll
    for (int j = 0; j < 3; ++j) {
        const FF ax = ff_two_sum(at(vs[j], kx), -ox), ay = ff_two_sum(at(vs[j], ky), -oy);
        const FF az = ff_two_sum(at(vs[j], kz), -oz);
        const FF px = ff_mul(sx, az), py = ff_mul(sy, az);
        h.az[j] = az;
        X[j] = ff_add(ax, ff_neg(px));
        Y[j] = ff_add(ay, ff_neg(py));
        eX[j] = kStep * (fabsf(ax.hi) + fabsf(px.hi)) + kFloor;
        eY[j] = kStep * (fabsf(ay.hi) + fabsf(py.hi)) + kFloor;
    }
    // w0 = cx*by - cy*bx, w1 = ax*cy - ay*cx, w2 = bx*ay - by*ax
    const int i0[3] = {2, 0, 1}, i1[3] = {1, 2, 0};
    FF W[3];
    float eW[3], mW = 0.0f, sW = 0.0f;
#pragma unroll
    for (int k = 0; k < 3; ++k) {
        const int m = i0[k], n = i1[k];
        const FF p0 = ff_mul(X[m], Y[n]), p1 = ff_mul(Y[m], X[n]);
        W[k] = ff_add(p0, ff_neg(p1));
        eW[k] = ((fabsf(X[m].hi) + eX[m]) * eY[n] + fabsf(Y[n].hi) * eX[m]) +
                ((fabsf(Y[m].hi) + eY[m]) * eX[n] + fabsf(X[n].hi) * eY[m]);
        eW[k] = 1.01f * eW[k] + kStep * (fabsf(p0.hi) + fabsf(p1.hi)) + kFloor;
        sW += eW[k];
        mW += fabsf(W[k].hi);
    }
    const FF det = ff_add(ff_add(W[0], W[1]), W[2]);
    const float edet = 1.01f * sW + 2.0f * kStep * mW + kFloor;
    const float adet = fabsf(det.hi);
    if (!(adet > 4.0f * edet) || !(adet >= 0x1p-100f) || !(adet < 0x1p100f)) return false;
#pragma unroll
    for (int k = 0; k < 3; ++k) {
        h.W[k] = W[k];
        h.eW[k] = eW[k];
    }
    h.det = det;
    h.edet = edet;
    h.u = ff_div2(W[1], det);
    h.v = ff_div2(W[2], det);
    const float au = fabsf(h.u.hi), av = fabsf(h.v.hi);
    h.eu = 1.01f * (eW[1] + 1.01f * au * edet) / (adet - edet) + kStep * au + kFloor;
    h.ev = 1.01f * (eW[2] + 1.01f * av * edet) / (adet - edet) + kStep * av + kFloor;
    h.w = ff_add(ff_add(FF{1.0f, 0.0f}, ff_neg(h.u)), ff_neg(h.v));
    h.ew = kStep * ((1.0f + au) + av) + kFloor;
    return finite1(h.u.hi) && finite1(h.v.hi) && finite1(h.w.hi) && h.ew < 0x1p-20f;
}

// The hit's t = (w0 (sz az) + w1 (sz bz) + w2 (sz cz)) / det, sz = 1/d_kz, as
// a float pair from hit_uv_ff's pairs, with a bound `et` on its distance from
// the reference's binary64 t (same step model).
CR_HDI FF hit_t_pair(const RayQuery& q, const HitUvFF& h, float& et) {
    constexpr float kStep = 0x1p-45f, kFloor = 0x1p-110f;
    const FF sz = ff_div2(FF{1.0f, 0.0f}, FF{at(q.d, q.kz), 0.0f});
    FF num{0.0f, 0.0f};
    float enum_ = 0.0f, mnum = 0.0f;
#pragma unroll
    for (int k = 0; k < 3; ++k) {
        const FF z = ff_mul(sz, h.az[k]);
        const float ez = 2.0f * kStep * fabsf(z.hi) + kFloor;
        const FF term = ff_mul(h.W[k], z);
        enum_ += ((fabsf(h.W[k].hi) + h.eW[k]) * ez + fabsf(z.hi) * h.eW[k]) * 1.01f + kStep * fabsf(term.hi);
        mnum += fabsf(term.hi);
        num = k == 0 ? term : ff_add(num, term);
    }
    enum_ += 2.0f * kStep * mnum + kFloor;
    const FF t = ff_div2(num, h.det);
    const float adet = fabsf(h.det.hi), at_ = fabsf(t.hi);
    et = 1.01f * (enum_ + 1.01f * at_ * h.edet) / (adet - h.edet) + kStep * at_ + kFloor;
    return t;
}

// float(t) where the pair's bound decides the reference's rounding.
CR_HDI bool hit_t_ff(const RayQuery& q, const HitUvFF& h, float& tf) {
    float et;
    const FF t = hit_t_pair(q, h, et);
    const float at_ = fabsf(t.hi);
    const float half_ulp = 0.5f * fminf(nextafterf(at_, INFINITY) - at_, at_ - nextafterf(at_, 0.0f));
    if (!(at_ >= 0x1p-100f) || !(at_ < 0x1p100f) || !(half_ulp - fabsf(t.lo) > 1.01f * et)) return false;
    tf = t.hi;
    return true;
}

enum FastResult { kReject = 0, kAccept = 1, kUnsure = 2 };

// FP32 watertight test with a forward error bound.  On kAccept, [tlo, thi]
// encloses the reference's binary64 t and that t lies certainly in [tmin, tmax].
//
// Bound derivation (u = 2^-24; exact quantities are the real-number values of
// the reference's formulas on the same float inputs; the reference's own
// binary64 rounding is ~2^-29 of every term below and is covered by the
// safety factors):
//   A = V - O            |Af - A|     <= 1.01u|Af| + |olow|        per component
//   x = A_kx - s*A_kz    |xf - x|     <= E_v = 1.02u(|A_kx|+|A_ky|+3|A_kz|+|x|+|y|) + 2|olow|_inf
//   w = cx*by - cy*bx    |wf - w|     <= (|cx|+|cy|)E_b + (|bx|+|by|)E_c + 2E_bE_c + 1.01u(|cx by|+|cy bx|+|wf|)
//   det = w0+w1+w2       |detf - det| <= sum E_w + 2.02u sum|w|
//   z_v = sz*A_v,kz      |zf - z|     <= 1.01|sz| eA_kz + 2.02u|zf|
//   N = sum w_v z_v      |Nf - N|     <= sum(|w|E_z + |z|E_w + E_wE_z) + 3.03u sum|w z|
//   t = N/det            |tf - t|     <= (E_N + |tf|E_det)/(|detf| - E_det) + 1.01u|tf|
// A sign is certain when |wf| exceeds its bound; the decision then equals the
// reference's because the binary64 value lies within the same interval.
CR_HDI int fast_triangle(const RayQuery& q, V3 va, V3 vb, V3 vc, float& tlo, float& thi) {
    const V3 a = va - q.o, b = vb - q.o, c = vc - q.o;
    const float az = at(a, q.kz), bz = at(b, q.kz), cz = at(c, q.kz);
    const float ax = at(a, q.kx) - q.sx * az, ay = at(a, q.ky) - q.sy * az;
    const float bx = at(b, q.kx) - q.sx * bz, by = at(b, q.ky) - q.sy * bz;
    const float cx = at(c, q.kx) - q.sx * cz, cy = at(c, q.ky) - q.sy * cz;
    const float w0 = cx * by - cy * bx;
    const float w1 = ax * cy - ay * cx;
    const float w2 = bx * ay - by * ax;
    const float k1 = 1.02f * kU;
    const float ea = k1 * ((fabsf(at(a, q.kx)) + fabsf(at(a, q.ky))) + 3.0f * fabsf(az) + (fabsf(ax) + fabsf(ay))) + q.ovec;
    const float eb = k1 * ((fabsf(at(b, q.kx)) + fabsf(at(b, q.ky))) + 3.0f * fabsf(bz) + (fabsf(bx) + fabsf(by))) + q.ovec;
    const float ec = k1 * ((fabsf(at(c, q.kx)) + fabsf(at(c, q.ky))) + 3.0f * fabsf(cz) + (fabsf(cx) + fabsf(cy))) + q.ovec;
    const float ma = fabsf(ax) + fabsf(ay), mb = fabsf(bx) + fabsf(by), mc = fabsf(cx) + fabsf(cy);
    const float k2 = 1.01f * kU;
    const float e0 = (mc * eb + mb * ec + 2.0f * eb * ec + k2 * (fabsf(cx * by) + fabsf(cy * bx) + fabsf(w0))) * 1.02f;
    const float e1 = (ma * ec + mc * ea + 2.0f * ec * ea + k2 * (fabsf(ax * cy) + fabsf(ay * cx) + fabsf(w1))) * 1.02f;
    const float e2 = (mb * ea + ma * eb + 2.0f * ea * eb + k2 * (fabsf(bx * ay) + fabsf(by * ax) + fabsf(w2))) * 1.02f;
    if (!finite1(e0) || !finite1(e1) || !finite1(e2) || !finite1(w0) || !finite1(w1) || !finite1(w2)) return kUnsure;
    const bool p0 = w0 > e0, p1 = w1 > e1, p2 = w2 > e2;
    const bool n0 = w0 < -e0, n1 = w1 < -e1, n2 = w2 < -e2;
    if ((p0 || p1 || p2) && (n0 || n1 || n2)) return kReject;  // certainly mixed signs
    if (!((p0 && p1 && p2) || (n0 && n1 && n2))) return kUnsure;
    const float wsum = fabsf(w0) + fabsf(w1) + fabsf(w2);
    const float det = (w0 + w1) + w2;
    const float edet = (e0 + e1) + e2 + 2.02f * kU * wsum;
    const float adet = fabsf(det);
    if (!(adet > 2.0f * edet)) return kUnsure;
    const float za = q.sz * az, zb = q.sz * bz, zc = q.sz * cz;
    const float num = (w0 * za + w1 * zb) + w2 * zc;
    const float asz = fabsf(q.sz) * 1.01f;
    const float eza = asz * (k2 * fabsf(az) + q.ovec) + 2.02f * kU * fabsf(za);
    const float ezb = asz * (k2 * fabsf(bz) + q.ovec) + 2.02f * kU * fabsf(zb);
    const float ezc = asz * (k2 * fabsf(cz) + q.ovec) + 2.02f * kU * fabsf(zc);
    const float enu = ((fabsf(w0) * eza + fabsf(za) * e0 + e0 * eza) + (fabsf(w1) * ezb + fabsf(zb) * e1 + e1 * ezb) +
                       (fabsf(w2) * ezc + fabsf(zc) * e2 + e2 * ezc)) +
                      3.03f * kU * ((fabsf(w0 * za) + fabsf(w1 * zb)) + fabsf(w2 * zc));
    const float t = num / det;
    if (!finite1(t) || !finite1(enu)) return kUnsure;
    const float et = ((enu + fabsf(t) * edet * 1.01f) / (adet - edet) + 1.01f * kU * fabsf(t)) * 1.05f + 1e-37f;
    tlo = t - et;
    thi = t + et;
    if (!finite1(thi) || !finite1(tlo)) return kUnsure;
    if (tlo > q.tmax || thi < q.tmin) return kReject;
    if (tlo < q.tmin || thi > q.tmax) return kUnsure;
    return kAccept;
}

// Alpha-mask decision of a certain FP32 hit without binary64: kAccept or
// kReject when the MASK test (reference src/geometry.cpp:279-292) has that
// outcome at every point the binary64 barycentrics could lie at, kUnsure
// otherwise.  The barycentrics come from the FP32 edge functions (bounds as in
// fast_triangle), their error is carried through the UV interpolation, the slot
// transform and the wrap into texel space; every texel the bilinear (or
// nearest) footprint can reach is read, and the test must hold over the range
// of alpha values those texels allow (narrowed for bilinear filtering by the
// blend's Lipschitz bound) with a relative margin of 2^-16, which covers the
// rounding of the filter's float blend.  Every bound is doubled.
CR_HDI int mask_fp32(const DevScene& sc, const RayQuery& q, uint32_t prim, uint32_t material, V3 va, V3 vb, V3 vc) {
    const DevMaterial& m = sc.materials[material];
    const DevSlot& slot = m.base_tex;
    if (slot.texture < 0) return kUnsure;
    const V3 a = va - q.o, b = vb - q.o, c = vc - q.o;
    const float az = at(a, q.kz), bz = at(b, q.kz), cz = at(c, q.kz);
    const float ax = at(a, q.kx) - q.sx * az, ay = at(a, q.ky) - q.sy * az;
    const float bx = at(b, q.kx) - q.sx * bz, by = at(b, q.ky) - q.sy * bz;
    const float cx = at(c, q.kx) - q.sx * cz, cy = at(c, q.ky) - q.sy * cz;
    const float w0 = cx * by - cy * bx;
    const float w1 = ax * cy - ay * cx;
    const float w2 = bx * ay - by * ax;
    const float k1 = 1.02f * kU;
    const float ea = k1 * ((fabsf(at(a, q.kx)) + fabsf(at(a, q.ky))) + 3.0f * fabsf(az) + (fabsf(ax) + fabsf(ay))) + q.ovec;
    const float eb = k1 * ((fabsf(at(b, q.kx)) + fabsf(at(b, q.ky))) + 3.0f * fabsf(bz) + (fabsf(bx) + fabsf(by))) + q.ovec;
    const float ec = k1 * ((fabsf(at(c, q.kx)) + fabsf(at(c, q.ky))) + 3.0f * fabsf(cz) + (fabsf(cx) + fabsf(cy))) + q.ovec;
    const float ma = fabsf(ax) + fabsf(ay), mb = fabsf(bx) + fabsf(by), mc = fabsf(cx) + fabsf(cy);
    const float k2 = 1.01f * kU;
    const float e0 = (mc * eb + mb * ec + 2.0f * eb * ec + k2 * (fabsf(cx * by) + fabsf(cy * bx) + fabsf(w0))) * 1.02f;
    const float e1 = (ma * ec + mc * ea + 2.0f * ec * ea + k2 * (fabsf(ax * cy) + fabsf(ay * cx) + fabsf(w1))) * 1.02f;
    const float e2 = (mb * ea + ma * eb + 2.0f * ea * eb + k2 * (fabsf(bx * ay) + fabsf(by * ax) + fabsf(w2))) * 1.02f;
    const float det = (w0 + w1) + w2;
    const float edet = (e0 + e1) + e2 + 2.02f * kU * (fabsf(w0) + fabsf(w1) + fabsf(w2));
    const float adet = fabsf(det);
    if (!(adet > 4.0f * edet)) return kUnsure;
    const float den = adet - edet;
    // Barycentrics of b (u), c (v) and a (1 - u - v) with their error bounds.
    const float bu = w1 / det, bv = w2 / det, bw = w0 / det;
    const float eu = 2.0f * ((e1 + fabsf(bu) * edet) / den + 2.0f * kU * fabsf(bu));
    const float ev = 2.0f * ((e2 + fabsf(bv) * edet) / den + 2.0f * kU * fabsf(bv));
    const float ew = 2.0f * ((e0 + fabsf(bw) * edet) / den + 2.0f * kU * fabsf(bw));
    // interpolateUv of the slot's set, rounded to float by the reference.
    const int off = slot.texcoord == 0 ? 0 : 2;
    const float* ua = &sc.uvs[3 * prim].x + off;
    const float* ub = &sc.uvs[3 * prim + 1].x + off;
    const float* uc = &sc.uvs[3 * prim + 2].x + off;
    float uv[2], euv[2];
    for (int d = 0; d < 2; ++d) {
        const float pa = bw * ua[d], pb = bu * ub[d], pc = bv * uc[d];
        uv[d] = (pa + pb) + pc;
        euv[d] = 2.0f * (ew * fabsf(ua[d]) + eu * fabsf(ub[d]) + ev * fabsf(uc[d]) +
                         4.0f * kU * (fabsf(pa) + fabsf(pb) + fabsf(pc)) + kU * fabsf(uv[d]));
    }
    // Slot transform (scale, rotation, offset) and the wrap, in float.
    const float cr = sf::to_float(sf::f64{slot.cos_bits}), sr = sf::to_float(sf::f64{slot.sin_bits});
    const float x = uv[0] * slot.scale[0], y = uv[1] * slot.scale[1];
    const float ex = euv[0] * fabsf(slot.scale[0]) + kU * fabsf(x), ey = euv[1] * fabsf(slot.scale[1]) + kU * fabsf(y);
    float tc[2], et[2];
    tc[0] = (cr * x - sr * y) + slot.offset[0];
    tc[1] = (sr * x + cr * y) + slot.offset[1];
    et[0] = 2.0f * (fabsf(cr) * ex + fabsf(sr) * ey + 4.0f * kU * (fabsf(cr * x) + fabsf(sr * y) + fabsf(slot.offset[0])) +
                    2.0f * kU * (fabsf(x) + fabsf(y)));
    et[1] = 2.0f * (fabsf(sr) * ex + fabsf(cr) * ey + 4.0f * kU * (fabsf(sr * x) + fabsf(cr * y) + fabsf(slot.offset[1])) +
                    2.0f * kU * (fabsf(x) + fabsf(y)));
    const DevTexture tex = sc.textures[slot.texture];
    const DevImage im = sc.images[tex.image];
    const int wrap[2] = {tex.wrap_s, tex.wrap_t};
    const int size[2] = {im.width, im.height};
    int64_t lo[2], hi[2];
    float pos[2], err[2];
    for (int d = 0; d < 2; ++d) {
        if (!finite1(tc[d]) || !finite1(et[d]) || fabsf(tc[d]) > 1048576.0f) return kUnsure;
        float t = tc[d];
        if (wrap[d] == 1) t = t < 0.0f ? 0.0f : (t > 1.0f ? 1.0f : t);
        else if (wrap[d] == 0) t = t - floorf(t);
        else {
            float r = fmodf(t, 2.0f);
            if (r < 0.0f) r += 2.0f;
            t = r <= 1.0f ? r : 2.0f - r;
        }
        const float X = t * float(size[d]) - (tex.linear ? 0.5f : 0.0f);
        const float eX = 2.0f * ((et[d] + 4.0f * kU) * float(size[d]) + 2.0f * kU * fabsf(X)) + 1e-4f;
        if (!(eX < 0.75f)) return kUnsure;
        lo[d] = int64_t(floorf(X - eX));
        hi[d] = int64_t(floorf(X + eX)) + (tex.linear ? 1 : 0);
        pos[d] = X;
        err[d] = eX;
    }
    auto alpha_at = [&](int64_t ix, int64_t iy) {
        const int tx = wrap_index(ix, im.width, tex.wrap_s), ty = wrap_index(iy, im.height, tex.wrap_t);
        return sc.texels[im.offset + uint64_t(ty) * uint64_t(im.width) + uint64_t(tx)].w;
    };
    float amin = 3.4e38f, amax = -3.4e38f;
    for (int64_t iy = lo[1]; iy <= hi[1]; ++iy)
        for (int64_t ix = lo[0]; ix <= hi[0]; ++ix) {
            const float al = alpha_at(ix, iy);
            amin = fminf(amin, al);
            amax = fmaxf(amax, al);
        }
    // Bilinear: the blend at the FP32 position, off from the blend at the
    // binary64 position by at most the footprint's alpha range times the
    // position error per axis (the blend is Lipschitz on the wrapped lattice).
    if (tex.linear) {
        const float fx0 = floorf(pos[0]), fy0 = floorf(pos[1]);
        const int64_t x0 = int64_t(fx0), y0 = int64_t(fy0);
        const float fx = pos[0] - fx0, fy = pos[1] - fy0;
        const float top = alpha_at(x0, y0) * (1 - fx) + alpha_at(x0 + 1, y0) * fx;
        const float bot = alpha_at(x0, y0 + 1) * (1 - fx) + alpha_at(x0 + 1, y0 + 1) * fx;
        const float est = top * (1 - fy) + bot * fy;
        const float e = (amax - amin) * (err[0] + err[1]) * 1.01f + 8.0f * kU * fmaxf(fabsf(amin), fabsf(amax));
        amin = fmaxf(amin, est - e);
        amax = fminf(amax, est + e);
    }
    const float f = m.base[3];
    if (f * amin * (1.0f - 0x1p-16f) >= m.alpha_cutoff && f * amin >= 0.0f) return kAccept;
    if (f * amax * (1.0f + 0x1p-16f) < m.alpha_cutoff && f * amax >= 0.0f) return kReject;
    return kUnsure;
}

// accepts(): single-sided opaque culling on the binary64 sign of dot(Ng, d)
// (reference src/geometry.cpp:279-292).
CR_HDI bool prim_accepts(const DevScene& s, uint32_t prim, uint32_t flags, V3 d) {
    if (flags & kPrimNeverHit) return false;
    if (flags & kPrimCull) {
        const CR_GLOBAL PrimAuxNg& n = gref(reinterpret_cast<const PrimAuxNg*>(s.prim_ng + prim));
        if (dot_sign_exact(V3{n.x, n.y, n.z}, d) >= 0) return false;
    }
    return true;
}

// Slab test against the children of a W-wide quantized node.  Each child's
// planes are decoded exactly as the builder evaluated them (layout.h QNode),
// giving a box that contains the child's float box, and then tested like a
// float box: FP32 slabs widened by 4 ulp relative plus the split-origin slack,
// so the test can only over-include.  Returns the hit mask and the widened
// entry distances.
// A child slot is present when its ref is not kEmptyChild (the builder and
// the cache-file check keep that equal to slot < child count).  Testing the
// refs makes the mask depend on them, so their load is issued with the
// planes' instead of after the mask test (a second round trip to memory on
// every visit that hits a child).
template <int W>
CR_HDI uint32_t node_hit(const RayQuery& q, const CR_GLOBAL QNode<W>& n, float limit, float near_out[W]) {
    const uint32_t ex = n.exps;
    const float sx = qscale(ex & 255u), sy = qscale((ex >> 8) & 255u), sz = qscale((ex >> 16) & 255u);
    const float px = n.px, py = n.py, pz = n.pz;
    const bool dx = !(q.inv[0] < 0), dy = !(q.inv[1] < 0), dz = !(q.inv[2] < 0);
    // Near and far plane words of each axis, picked once per word by the ray's
    // direction sign (a positive direction enters at the low plane); each
    // child's byte then converts straight to float.
    uint32_t nw[3][W / 4], fw[3][W / 4];
#pragma unroll
    for (int w = 0; w < W / 4; ++w) {
        nw[0][w] = dx ? n.qlo[0][w] : n.qhi[0][w];
        fw[0][w] = dx ? n.qhi[0][w] : n.qlo[0][w];
        nw[1][w] = dy ? n.qlo[1][w] : n.qhi[1][w];
        fw[1][w] = dy ? n.qhi[1][w] : n.qlo[1][w];
        nw[2][w] = dz ? n.qlo[2][w] : n.qhi[2][w];
        fw[2][w] = dz ? n.qhi[2][w] : n.qlo[2][w];
    }
    uint32_t mask = 0;
#pragma unroll
    for (int c = 0; c < W; ++c) {
        const int w = c / 4, sh = 8 * (c % 4);
        const float nx = (qdecode(px, (nw[0][w] >> sh) & 255u, sx) - q.o.x) * q.inv[0];
        const float fx = (qdecode(px, (fw[0][w] >> sh) & 255u, sx) - q.o.x) * q.inv[0];
        const float ny = (qdecode(py, (nw[1][w] >> sh) & 255u, sy) - q.o.y) * q.inv[1];
        const float fy = (qdecode(py, (fw[1][w] >> sh) & 255u, sy) - q.o.y) * q.inv[1];
        const float nz = (qdecode(pz, (nw[2][w] >> sh) & 255u, sz) - q.o.z) * q.inv[2];
        const float fz = (qdecode(pz, (fw[2][w] >> sh) & 255u, sz) - q.o.z) * q.inv[2];
        float tn = fmaxf(fmaxf(q.tmin, nx), fmaxf(ny, nz));
        float tf = fminf(fminf(limit, fx), fminf(fy, fz));
        // Widening 4u|t| + slack: the product by 2^-22 is exact, so the fused
        // form rounds like the separate multiply and add.
        tn = tn - __builtin_fmaf(fabsf(tn), 4.0f * kU, q.slack);
        tf = tf + __builtin_fmaf(fabsf(tf), 4.0f * kU, q.slack);
        near_out[c] = tn;
        if (tn <= tf && n.child[c] != kEmptyChild) mask |= 1u << c;
    }
    return mask;
}

// One comparator of a sorting network over (entry distance, ref) pairs:
// ascending distances.  A hit child's distance is finite (an entry at +inf
// widens to NaN and fails the test), a missed child's is +inf.
CR_HDI void sort2(float& ka, uint32_t& ra, float& kb, uint32_t& rb) {
    const bool swap = kb < ka;
    const float k0 = swap ? kb : ka, k1 = swap ? ka : kb;
    const uint32_t r0 = swap ? rb : ra, r1 = swap ? ra : rb;
    ka = k0;
    kb = k1;
    ra = r0;
    rb = r1;
}

// Ascending sort of W keys with their refs (W = 4: 5 comparators; W = 8: the
// 19-comparator network).
template <int W>
CR_HDI void sort_children(float (&k)[W], uint32_t (&r)[W]) {
    if constexpr (W == 4) {
        sort2(k[0], r[0], k[1], r[1]);
        sort2(k[2], r[2], k[3], r[3]);
        sort2(k[0], r[0], k[2], r[2]);
        sort2(k[1], r[1], k[3], r[3]);
        sort2(k[1], r[1], k[2], r[2]);
    } else {
        constexpr int net[19][2] = {{0, 2}, {1, 3}, {4, 6}, {5, 7}, {0, 4}, {1, 5}, {2, 6}, {3, 7}, {0, 1}, {2, 3},
                                    {4, 5}, {6, 7}, {2, 4}, {3, 5}, {1, 4}, {3, 6}, {1, 2}, {3, 4}, {5, 6}};
#pragma unroll
        for (int i = 0; i < 19; ++i) sort2(k[net[i][0]], r[net[i][0]], k[net[i][1]], r[net[i][1]]);
    }
}

constexpr int kStackSize = 160;

struct StackEntry {
    uint32_t ref;
    float near;
};

struct ClosestResult {
    uint32_t prim;  // leaf-order primitive, kNoTriangle on miss
    uint32_t id;    // finalized triangle index
    bool exact;       // t,u,v hold the binary64 values
    bool overflow;    // traversal stack exhausted (reported as an error)
    bool tex_error;   // alpha-mask texture coordinates not finite (reported as an error)
    bool need_exact;  // FP32-only traversal left candidates (see exact_work)
    sf::f64 t, u, v;
    float tlo, thi;  // FP32 enclosure of the binary64 t (or the query limit)
    uint32_t ncand;  // candidates met by the FP32-only traversal (kCand + 1: too many)
    // The first kCand candidates, in storage the caller owns (the FP32-only
    // traversal needs it; null elsewhere).  Candidates are rare: held here
    // they would cost registers in every traversal loop, and as arrays
    // indexed at run time they would move the whole result to private memory.
    Pending* cands;
};

// Replace the current best by an exactly evaluated candidate when it wins the
// reference's (t, id) ordering.
CR_HDI void take_exact(ClosestResult& best, uint32_t prim, uint32_t id, const ExactHit& e) {
    best.prim = prim;
    best.id = id;
    best.exact = true;
    best.t = e.t;
    best.u = e.u;
    best.v = e.v;
    const float tf = sf::to_float(e.t);
    best.tlo = next_down(tf);
    best.thi = next_up(tf);
}

// FP32-only traversal: a primitive it cannot decide alone becomes a candidate
// for the exact pass; the traversal goes on.
CR_HDI void add_cand(ClosestResult& b, uint32_t prim, float lo) {
    if (b.ncand < kCand) {
        b.cands->cand[b.ncand] = prim;
        b.cands->lo[b.ncand] = lo;
    }
    if (b.ncand <= kCand) ++b.ncand;
    b.need_exact = true;
}

// FP32-only decision for one tested primitive (fr != kReject, [clo, chi] the
// FP32 enclosure of its t when fast_triangle got that far, else 0).  Certain
// hits update the best, whose upper bound is the traversal's limit; a certain
// hit whose order against the best is undecided keeps the smaller upper bound
// as the best and makes the other a candidate; undecided tests and alpha-mask
// primitives become candidates.  Returns true when an any-hit query found a
// certain blocker.
// MASK=false leaves every alpha-mask primitive to the exact pass (same
// results; the traversal loop then carries no texture code, for scenes
// without alpha masks).
template <bool ANY, bool MASK = true>
CR_HDI bool fp32_take(const DevScene& s, const RayQuery& q, ClosestResult& best, uint32_t i, uint32_t id,
                      uint32_t flags, int fr, float clo, float chi) {
    if (!ANY && clo > best.thi) return false;           // certainly beyond the limit (also when undecided)
    if (!ANY && id == best.id && best.prim != kNoTriangle) return false;  // another reference of the best's triangle
    if (!prim_accepts(s, i, flags, q.d)) return false;  // culling does not depend on t
    bool certain = fr == kAccept && !(flags & kPrimMaskTexture);
    if (MASK && fr == kAccept && !certain) {  // alpha mask: decided in FP32 where the texels allow
        const CR_GLOBAL PrimRec& p = gref(s.prims + i);
        const int mr = mask_fp32(s, q, i, p.material, V3{p.ax, p.ay, p.az}, V3{p.bx, p.by, p.bz}, V3{p.cx, p.cy, p.cz});
        if (mr == kReject) return false;
        certain = mr == kAccept;
    }
    if (certain) {
        if (ANY || best.prim == kNoTriangle || chi < best.tlo) {  // certainly nearer than the best
            best.prim = i;
            best.id = id;
            best.exact = false;
            best.tlo = clo;
            best.thi = chi;
            return ANY;
        }
        if (chi < best.thi) {
            add_cand(best, best.prim, best.tlo);
            best.prim = i;
            best.id = id;
            best.tlo = clo;
            best.thi = chi;
        } else {
            add_cand(best, i, clo);
        }
        return false;
    }
    add_cand(best, i, clo);
    return false;
}

// What an FP32-only result still needs: 0 none, 1 exact resolution of its
// candidates (those that may lie within the best's bound), 2 the exact
// traversal (more candidates than were kept).
template <bool ANY>
CR_HDI int exact_work(const ClosestResult& b) {
    if (b.overflow || b.tex_error || b.ncand == 0) return 0;
    if (ANY && b.prim != kNoTriangle) return 0;  // a certain blocker
    if (b.ncand > kCand) return 2;
    if (ANY) return 1;
    for (uint32_t k = 0; k < b.ncand; ++k)
        if (!(b.cands->lo[k] > b.thi)) return 1;
    return 0;
}

// The exact pass's record of an FP32-only result with work != 0 (its
// candidates then sit in b.cands).  Unused candidate slots are zero.
CR_HDI Pending pending_of(const ClosestResult& b, int work) {
    Pending p;
    p.best = b.prim;
    p.n = b.ncand > kCand ? kCand : b.ncand;
    p.thi = b.thi;
    p.full = work == 2 ? 1u : 0u;
    for (uint32_t k = 0; k < kCand; ++k) {
        p.cand[k] = k < p.n ? b.cands->cand[k] : 0u;
        p.lo[k] = k < p.n ? b.cands->lo[k] : 0.0f;
    }
    return p;
}

// Per-query work counters (wide nodes visited, triangles tested, exact fallbacks).
struct TraceCount {
    uint32_t nodes, tris, fallbacks;
};

// Tests a leaf's primitives in the exact traversal: FP32 first, the binary64
// emulation for what FP32 leaves undecided.  Returns true when traversal can
// stop (any-hit found).  The FP32-only traversal uses leaf_fp32 instead.
template <bool ANY, bool MASK = true>
CR_HD inline bool leaf_prims(const DevScene& s, const PrimRec* prims, RayQuery& q, uint32_t first, uint32_t count,
                             ClosestResult& best, TraceCount& tc) {
    for (uint32_t i = first; i < first + count; ++i) {
        const CR_GLOBAL PrimRec& p = gref(prims + i);
        if (p.id == q.ignore) continue;
        const V3 va{p.ax, p.ay, p.az}, vb{p.bx, p.by, p.bz}, vc{p.cx, p.cy, p.cz};
        ++tc.tris;
        float clo = 0, chi = 0;
        const int fr = fast_triangle(q, va, vb, vc, clo, chi);
        if (fr == kReject) continue;
        if (fr == kAccept && !(p.flags & kPrimMaskTexture)) {
            if (!ANY && clo > best.thi) continue;  // certainly beyond the current limit
            if (!prim_accepts(s, i, p.flags, q.d)) continue;
            if (ANY) {
                best.prim = i;
                best.id = p.id;
                best.exact = false;
                best.tlo = clo;
                best.thi = chi;
                return true;
            }
            if (chi < best.tlo) {  // certainly nearer than the current best
                best.prim = i;
                best.id = p.id;
                best.exact = false;
                best.tlo = clo;
                best.thi = chi;
                continue;
            }
        }
        if (fr == kAccept && !ANY && clo > best.thi) continue;
        // Undecided in FP32, or an alpha mask that needs the binary64
        // barycentrics: decide with the binary64 emulation.
        ++tc.fallbacks;
        if (!ANY && !best.exact) {
            const CR_GLOBAL PrimRec& bp = gref(prims + best.prim);
            const ExactHit e = exact_triangle(q, V3{bp.ax, bp.ay, bp.az}, V3{bp.bx, bp.by, bp.bz},
                                              V3{bp.cx, bp.cy, bp.cz}, sf::inf(0));
            take_exact(best, best.prim, best.id, e);
        }
        const ExactHit e = exact_triangle(q, va, vb, vc, ANY ? sf::inf(0) : best.t);
        if (!e.hit) continue;
        if (!prim_accepts(s, i, p.flags, q.d)) continue;
        if (p.flags & kPrimMaskTexture) {
            bool bad = false;
            const bool keep = mask_accepts(s, i, p.material, e.u, e.v, bad);
            if (bad) {
                best.tex_error = true;
                return true;
            }
            if (!keep) continue;
        }
        if (ANY) {
            take_exact(best, i, p.id, e);
            return true;
        }
        if (best.prim == kNoTriangle || sf::lt(e.t, best.t) || (sf::eq(e.t, best.t) && p.id < best.id))
            take_exact(best, i, p.id, e);
    }
    return false;
}

// FP32-only leaf test.  Each primitive record is read whole before any of its
// fields decides anything: the ignored-id and never-hit checks come after the
// triangle test (both only turn its outcome into a rejection, and fp32_take
// rejects a never-hit primitive without side effects), so the record's id,
// vertices and flags are loaded together instead of one after another.
template <bool ANY, bool MASK = true>
CR_HD inline bool leaf_fp32(const DevScene& s, const PrimRec* prims, RayQuery& q, uint32_t first, uint32_t count,
                            ClosestResult& best, TraceCount& tc) {
    for (uint32_t i = first; i < first + count; ++i) {
        const CR_GLOBAL PrimRec& p = gref(prims + i);
        const uint32_t id = p.id, flags = p.flags;
        float clo = 0, chi = 0;
        int fr = fast_triangle(q, V3{p.ax, p.ay, p.az}, V3{p.bx, p.by, p.bz}, V3{p.cx, p.cy, p.cz}, clo, chi);
        tc.tris += id != q.ignore ? 1u : 0u;
        if (id == q.ignore || (flags & kPrimNeverHit)) fr = kReject;
        if (fr == kReject) continue;
        if (fp32_take<ANY, MASK>(s, q, best, i, id, flags, fr, clo, chi)) return true;
    }
    return false;
}

// FP32-only leaf test with the primitive records of each group of four loaded
// before any is tested, so their loads are in flight together.  Same decisions
// as leaf_fp32.
#ifndef CR_LEAF_PRELOAD
#define CR_LEAF_PRELOAD 0
#endif
#ifndef CR_LEAF_GROUP
#define CR_LEAF_GROUP 4
#endif
// Keep the scene's node and primitive array pointers of a query in registers
// (1) or read them from the scene record at each visit (0).
#ifndef CR_CACHE_PTRS
#define CR_CACHE_PTRS 1
#endif
template <bool ANY, bool MASK = true>
CR_HD inline bool leaf_fast(const DevScene& s, const PrimRec* prims, RayQuery& q, uint32_t first, uint32_t count,
                            ClosestResult& best, TraceCount& tc) {
    const uint32_t end = first + count;
    for (uint32_t base = first; base < end; base += CR_LEAF_GROUP) {
        const uint32_t m = end - base < CR_LEAF_GROUP ? end - base : CR_LEAF_GROUP;
        PrimRec pr[CR_LEAF_GROUP];
#pragma unroll
        for (uint32_t j = 0; j < CR_LEAF_GROUP; ++j)
            if (j < m) {
                const CR_GLOBAL PrimRec& g = gref(prims + base + j);
                pr[j].ax = g.ax, pr[j].ay = g.ay, pr[j].az = g.az, pr[j].id = g.id;
                pr[j].bx = g.bx, pr[j].by = g.by, pr[j].bz = g.bz, pr[j].flags = g.flags;
                pr[j].cx = g.cx, pr[j].cy = g.cy, pr[j].cz = g.cz, pr[j].material = g.material;
            }
#pragma unroll
        for (uint32_t j = 0; j < CR_LEAF_GROUP; ++j) {
            if (j >= m) break;
            const PrimRec& p = pr[j];
            if (p.id == q.ignore) continue;
            ++tc.tris;
            float clo = 0, chi = 0;
            const int fr = fast_triangle(q, V3{p.ax, p.ay, p.az}, V3{p.bx, p.by, p.bz}, V3{p.cx, p.cy, p.cz}, clo, chi);
            if (fr == kReject) continue;
            if (fp32_take<ANY, MASK>(s, q, best, base + j, p.id, p.flags, fr, clo, chi)) return true;
        }
    }
    return false;
}

// Bvh::intersect on the 4-wide tree (reference src/geometry.cpp:554-639) as a
// resumable traversal: step() performs one node or leaf visit, so a
// persistent kernel can interleave several queries per lane.  EXACT=false is
// the FP32-only fast path (see leaf_prims).
// Traversal stack = a small ring of the newest entries (in LDS on the device:
// `LdsRing`; private arrays otherwise: `LocalRing`) on top of a private
// overflow array holding the older entries.  One LIFO stack: when the ring is
// full, its oldest entry moves to the overflow array; pops take the ring
// first.  Entries are {child ref, entry distance}.
// The Traversal holds only pointers to the ring and overflow storage, which
// its caller declares: both are indexed at run time, and an object with such
// an array inside stays in private (scratch) memory as a whole, so the best
// hit, its bound and the stack counters would be stored and reloaded there at
// every visit.  With the arrays outside, that state stays in registers.
struct LocalRing {
    static constexpr int kSize = 8;
    uint32_t* ref;  // kSize entries each, owned by the caller
    float* near;
    CR_HDI void put(int i, uint32_t r, float n) {
        ref[i] = r;
        near[i] = n;
    }
    CR_HDI void get(int i, uint32_t& r, float& n) const {
        r = ref[i];
        n = near[i];
    }
};

template <bool ANY, bool EXACT = true, class Ring = LocalRing, bool PRELOAD = (CR_LEAF_PRELOAD != 0), bool MASK = true,
          int W = 4>
struct Traversal {
    ClosestResult best;
    StackEntry* stack;  // overflow (kStackSize entries, owned by the caller): entries older than the ring's
    Ring ring;
    Pending* cands;  // candidate storage of an FP32-only traversal (owned by the caller)
    int sp;    // entries in `stack`
    int head;  // ring slot of the next push
    int nr;    // entries in the ring
    uint32_t ref;
    uint32_t leaf;  // postponed leaf (kEmptyChild: none), see inner()/leaves()
    const QNode<W>* nodes;  // the scene's arrays, read once per query
    const PrimRec* prims;

    CR_HDI void push(uint32_t r, float n) {
        if (nr == Ring::kSize) {  // ring full: its oldest entry (at head) moves to the overflow array
            ring.get(head, stack[sp].ref, stack[sp].near);
            ++sp;
            --nr;
        }
        ring.put(head, r, n);
        head = head + 1 == Ring::kSize ? 0 : head + 1;
        ++nr;
    }

    CR_HDI bool pop(StackEntry& e) {
        if (nr > 0) {
            head = head == 0 ? Ring::kSize - 1 : head - 1;
            ring.get(head, e.ref, e.near);
            --nr;
            return true;
        }
        if (sp > 0) {
            e = stack[--sp];
            return true;
        }
        return false;
    }

    // Returns true when there is nothing to traverse (invalid query, empty tree).
    CR_HDI bool begin(const DevScene& s, const RayQuery& q) {
        best.prim = kNoTriangle;
        best.id = 0;
        best.exact = true;
        best.overflow = false;
        best.tex_error = false;
        best.need_exact = false;
        best.ncand = 0;
        best.cands = cands;
        best.t = sf::from_float(q.tmax);
        best.u = best.v = sf::zero(0);
        best.tlo = best.thi = q.tmax;
        sp = 0;
        head = 0;
        nr = 0;
        ref = 0;
        leaf = kEmptyChild;
        if constexpr (W == 8) nodes = s.nodes8;
        else nodes = s.nodes4;
        prims = s.prims;
        return !q.valid || s.node_count4 == 0 || q.tmax < q.tmin;
    }

    CR_HDI bool test_leaf(const DevScene& s, RayQuery& q, uint32_t r, TraceCount& tc) {
        const PrimRec* pr = CR_CACHE_PTRS ? prims : s.prims;
        if constexpr (!EXACT && PRELOAD) return leaf_fast<ANY, MASK>(s, pr, q, leaf_first(r), leaf_count(r), best, tc);
        else if constexpr (!EXACT) return leaf_fp32<ANY, MASK>(s, pr, q, leaf_first(r), leaf_count(r), best, tc);
        else return leaf_prims<ANY, MASK>(s, pr, q, leaf_first(r), leaf_count(r), best, tc);
    }

    // Visit of the inner node `ref`: pushes the farther hit children, moves to
    // the nearest one (or to the next stack entry).  Returns false when the
    // stack bound is exceeded (best.overflow set).
    CR_HDI bool visit_node(const DevScene& s, RayQuery& q, TraceCount& tc) {
        const CR_GLOBAL QNode<W>& n = gref(nodes + ref);
        ++tc.nodes;
        float nears[W];
        const uint32_t mask = node_hit<W>(q, n, best.thi, nears);
        if (!mask) {
            next_ref();
            return true;
        }
        const int hits = __builtin_popcount(mask);
        if (sp + nr + hits - 1 > kStackSize) {
            best.overflow = true;
            return false;
        }
        // Hit children sorted by entry distance (missed ones last): visit the
        // nearest, push the others farthest first so the next pop is the
        // second nearest.  The order only affects how soon the bound shrinks,
        // never the result.
        float key[W];
        uint32_t ch[W];
#pragma unroll
        for (int c = 0; c < W; ++c) {
            key[c] = ((mask >> c) & 1u) ? nears[c] : kInf;
            ch[c] = n.child[c];
        }
        sort_children<W>(key, ch);
#pragma unroll
        for (int k = W - 1; k >= 1; --k)
            if (k < hits) push(ch[k], key[k]);
        ref = ch[0];
        return true;
    }

    // Next stack entry that may still hold a nearer hit (kEmptyChild when none).
    CR_HDI void next_ref() {
        StackEntry e;
        while (pop(e)) {
            if (!ANY && e.near > best.thi) continue;
            ref = e.ref;
            return;
        }
        ref = kEmptyChild;
    }

    // Postponed-leaf traversal (Aila and Laine's speculative while-while
    // scheme): inner() visits inner nodes only; the first leaf reached is
    // parked in `leaf` and the walk goes on from the stack, a second one waits
    // in `ref`.  leaves() tests the parked leaf and a waiting one.  A wave
    // alternates the two phases so that its lanes load and test the same kind
    // of record at the same time.  Both return true when the query is
    // finished; the result does not depend on how the phases are interleaved.
    CR_HDI bool can_descend() const { return ref != kEmptyChild && !ref_is_leaf(ref); }
    CR_HDI bool has_leaf() const { return leaf != kEmptyChild || ref_is_leaf(ref); }

    CR_HDI bool inner(const DevScene& s, RayQuery& q, TraceCount& tc) {
        if (can_descend() && !visit_node(s, q, tc)) return true;
        if (ref_is_leaf(ref) && leaf == kEmptyChild) {
            leaf = ref;
            next_ref();
        }
        return ref == kEmptyChild && leaf == kEmptyChild;
    }

    CR_HDI bool leaves(const DevScene& s, RayQuery& q, TraceCount& tc) {
        if (leaf != kEmptyChild) {
            if (test_leaf(s, q, leaf, tc)) return true;
            leaf = kEmptyChild;
        }
        if (ref_is_leaf(ref)) {
            if (test_leaf(s, q, ref, tc)) return true;
            next_ref();
            if (ref_is_leaf(ref)) {
                leaf = ref;
                next_ref();
            }
        }
        return ref == kEmptyChild && leaf == kEmptyChild;
    }

    // One visit; returns true when the query is finished (result in best).
    CR_HDI bool step(const DevScene& s, RayQuery& q, TraceCount& tc) {
        if (!ref_is_leaf(ref)) {
            if (!visit_node(s, q, tc)) return true;
            return ref == kEmptyChild;
        }
        if (test_leaf(s, q, ref, tc)) return true;
        next_ref();
        return ref == kEmptyChild;
    }
};

// One whole traversal.  EXACT=false (FP32 only) needs `cands`, storage for
// the candidates its result refers to; the exact traversal keeps none.
template <bool ANY, bool EXACT = true, int W = 4>
CR_HD inline ClosestResult traverse(const DevScene& s, RayQuery& q, TraceCount& tc, Pending* cands = nullptr) {
    StackEntry overflow[kStackSize];
    uint32_t ring_ref[LocalRing::kSize];
    float ring_near[LocalRing::kSize];
    Traversal<ANY, EXACT, LocalRing, (CR_LEAF_PRELOAD != 0), true, W> tr;
    tr.stack = overflow;
    tr.ring = LocalRing{ring_ref, ring_near};
    tr.cands = cands;
    if (!tr.begin(s, q))
        while (!tr.step(s, q, tc)) {
        }
    return tr.best;
}

// Exact (t, u, v) of a known hit primitive.
CR_HDI ExactHit exact_of(const DevScene& s, RayQuery& q, uint32_t prim) {
    const PrimRec p = s.prims[prim];
    return exact_triangle(q, V3{p.ax, p.ay, p.az}, V3{p.bx, p.by, p.bz}, V3{p.cx, p.cy, p.cz}, sf::inf(0));
}

// Exact (u, v), and t when `want_t`, of the closest hit a traversal decided.
CR_HDI ExactHit exact_uv_of(const DevScene& s, RayQuery& q, uint32_t prim, bool want_t) {
    const PrimRec p = s.prims[prim];
    return exact_hit_uv(q, V3{p.ax, p.ay, p.az}, V3{p.bx, p.by, p.bz}, V3{p.cx, p.cy, p.cz}, want_t);
}

// Exact pass over the candidates an FP32-only traversal left (exact_work 1):
// the best so far and every candidate that may lie within its bound are
// evaluated in binary64 and ordered by (t, id) as the reference orders hits
// (reference src/geometry.cpp:609); any-hit stops at the first candidate that
// hits.  Every primitive the FP32 pass did not keep was decided there (certain
// miss, or certainly farther than a certain hit), so this equals the exact
// traversal's result.
template <bool ANY>
CR_HD inline ClosestResult resolve_candidates(const DevScene& s, RayQuery& q, const Pending& pd, TraceCount& tc) {
    ClosestResult best;
    best.prim = kNoTriangle;
    best.id = 0;
    best.exact = true;
    best.overflow = false;
    best.tex_error = false;
    best.need_exact = false;
    best.ncand = 0;
    best.cands = nullptr;
    best.t = sf::from_float(q.tmax);
    best.u = best.v = sf::zero(0);
    best.tlo = best.thi = q.tmax;
    if (!ANY && pd.best != kNoTriangle) {
        ++tc.fallbacks;
        take_exact(best, pd.best, s.prims[pd.best].id, exact_of(s, q, pd.best));
    }
    for (uint32_t k = 0; k < pd.n; ++k) {
        if (!ANY && pd.lo[k] > pd.thi) continue;
        const uint32_t i = pd.cand[k];
        const CR_GLOBAL PrimRec& p = gref(s.prims + i);
        ++tc.fallbacks;
        const ExactHit e = exact_triangle(q, V3{p.ax, p.ay, p.az}, V3{p.bx, p.by, p.bz}, V3{p.cx, p.cy, p.cz},
                                          ANY ? sf::inf(0) : best.t);
        if (!e.hit) continue;
        if (!prim_accepts(s, i, p.flags, q.d)) continue;
        if (p.flags & kPrimMaskTexture) {
            bool bad = false;
            const bool keep = mask_accepts(s, i, p.material, e.u, e.v, bad);
            if (bad) {
                best.tex_error = true;
                return best;
            }
            if (!keep) continue;
        }
        if (ANY) {
            take_exact(best, i, p.id, e);
            return best;
        }
        if (best.prim == kNoTriangle || sf::lt(e.t, best.t) || (sf::eq(e.t, best.t) && p.id < best.id))
            take_exact(best, i, p.id, e);
    }
    return best;
}

// The exact pass the kernels run after an FP32-only traversal that left work
// (exact_work != 0): candidate resolution, or the exact traversal.
template <bool ANY>
CR_HD inline ClosestResult exact_pass(const DevScene& s, RayQuery& q, const Pending& pd, TraceCount& tc) {
    if (pd.full) return traverse<ANY, true>(s, q, tc);
    return resolve_candidates<ANY>(s, q, pd, tc);
}

}  // namespace cr
