// vmath.h — float vector algebra and the per-sample random stream, host+device.
//
// Operation order follows the reference value types exactly (reference
// src/math.h:15-128) because per-sample parity requires identical float
// rounding: scalar division is multiplication by the reciprocal, dot products
// sum x, y, z left to right, luminance uses the Rec.709 weights, and every
// product/sum is a separate rounded operation (device code is compiled with
// -ffp-contract=off so no multiply-add is fused behind our back).
#pragma once
#include "sf64.h"

#include <math.h>
#include <stdint.h>

namespace cr {

constexpr float kPi = 3.14159265358979323846f;
constexpr float kInf = __builtin_huge_valf();
constexpr float kFltMax = 3.40282346638528859812e+38f;
constexpr float kFltEps = 1.1920928955078125e-7f;
constexpr uint32_t kNoTriangle = 0xffffffffu;

struct V3 {
    float x, y, z;
};

CR_HDI V3 v3(float s) { return {s, s, s}; }
CR_HDI V3 v3(float a, float b, float c) { return {a, b, c}; }
CR_HDI float at(V3 v, int i) { return i == 0 ? v.x : (i == 1 ? v.y : v.z); }
CR_HDI void set(V3& v, int i, float s) {
    if (i == 0) v.x = s;
    else if (i == 1) v.y = s;
    else v.z = s;
}
CR_HDI V3 operator+(V3 a, V3 b) { return {a.x + b.x, a.y + b.y, a.z + b.z}; }
CR_HDI V3 operator-(V3 a, V3 b) { return {a.x - b.x, a.y - b.y, a.z - b.z}; }
CR_HDI V3 operator-(V3 a) { return {-a.x, -a.y, -a.z}; }
CR_HDI V3 operator*(V3 a, V3 b) { return {a.x * b.x, a.y * b.y, a.z * b.z}; }
CR_HDI V3 operator*(V3 a, float b) { return {a.x * b, a.y * b, a.z * b}; }
CR_HDI V3 operator*(float a, V3 b) { return b * a; }
CR_HDI V3 operator/(V3 a, float b) { return a * (1.0f / b); }
CR_HDI float dot(V3 a, V3 b) { return a.x * b.x + a.y * b.y + a.z * b.z; }
CR_HDI V3 cross(V3 a, V3 b) { return {a.y * b.z - a.z * b.y, a.z * b.x - a.x * b.z, a.x * b.y - a.y * b.x}; }
CR_HDI float length_sq(V3 a) { return dot(a, a); }
CR_HDI float length(V3 a) { return sqrtf(length_sq(a)); }
CR_HDI V3 normalize(V3 a) {
    const float l = length(a);
    return l > 0 ? a / l : V3{0, 0, 0};
}
CR_HDI bool finite1(float x) { return (sf::fbits(x) & 0x7f800000u) != 0x7f800000u; }
CR_HDI bool finite(V3 a) { return finite1(a.x) && finite1(a.y) && finite1(a.z); }
CR_HDI bool nonzero(V3 a) { return a.x != 0 || a.y != 0 || a.z != 0; }
// std::min / std::max / std::clamp semantics (first argument wins ties and NaN).
CR_HDI float fmin_(float a, float b) { return b < a ? b : a; }
CR_HDI float fmax_(float a, float b) { return a < b ? b : a; }
CR_HDI float clampf(float v, float lo, float hi) { return v < lo ? lo : (hi < v ? hi : v); }
CR_HDI V3 clamp(V3 a, float lo, float hi) { return {clampf(a.x, lo, hi), clampf(a.y, lo, hi), clampf(a.z, lo, hi)}; }
CR_HDI float max_component(V3 a) {
    float m = a.x;
    if (m < a.y) m = a.y;
    if (m < a.z) m = a.z;
    return m;
}
CR_HDI float min_component(V3 a) {
    float m = a.x;
    if (a.y < m) m = a.y;
    if (a.z < m) m = a.z;
    return m;
}
CR_HDI float luminance(V3 a) { return dot(a, V3{0.2126f, 0.7152f, 0.0722f}); }
CR_HDI V3 mix(V3 a, V3 b, float t) { return a * (1 - t) + b * t; }
CR_HDI V3 reflect(V3 i, V3 n) { return i - 2 * dot(i, n) * n; }

// Orthonormal basis whose local +Z is the given unit normal.
struct Frame {
    V3 x, y, z;
};
CR_HDI Frame make_frame(V3 n) {
    Frame f;
    f.z = n;
    f.x = normalize(cross(fabsf(n.z) < 0.999f ? V3{0, 0, 1} : V3{0, 1, 0}, n));
    f.y = cross(n, f.x);
    return f;
}
CR_HDI V3 to_local(const Frame& f, V3 v) { return {dot(v, f.x), dot(v, f.y), dot(v, f.z)}; }
CR_HDI V3 to_world(const Frame& f, V3 v) { return f.x * v.x + f.y * v.y + f.z * v.z; }

// nextafterf toward +inf / -inf / zero for finite or infinite inputs.
CR_HDI float next_up(float x) {
    if (x != x || x == kInf) return x;
    if (x == 0) return sf::bitsf(1);
    const uint32_t u = sf::fbits(x);
    return sf::bitsf(x > 0 ? u + 1 : u - 1);
}
CR_HDI float next_down(float x) { return -next_up(-x); }
CR_HDI float next_toward_zero(float x) {
    if (x != x || x == 0) return x;
    return sf::bitsf(sf::fbits(x) - 1);
}

// SplitMix64 stream, one per (seed, pixel, sample) (reference src/math.h:106-121).
struct Rng {
    uint64_t state;
    CR_HDI uint32_t next() {
        state += 0x9e3779b97f4a7c15ull;
        uint64_t z = state;
        z = (z ^ (z >> 30)) * 0xbf58476d1ce4e5b9ull;
        z = (z ^ (z >> 27)) * 0x94d049bb133111ebull;
        return uint32_t((z ^ (z >> 31)) >> 32);
    }
    // 24 random bits -> float in [0,1); bits24() exposes the integer for
    // exact CDF inversion by threshold tables.
    CR_HDI uint32_t bits24() { return next() >> 8; }
    CR_HDI float uniform() { return float(bits24()) * (1.0f / 16777216.0f); }
};

CR_HDI uint64_t sample_seed(uint64_t seed, uint64_t pixel, uint64_t sample) {
    Rng h{seed ^ (pixel * 0xd2b74407b1ce6e93ull) ^ (sample * 0xca5a826395121157ull)};
    const uint64_t hi = h.next();
    return (hi << 32) | h.next();
}

// ---- binary64 helpers on float vectors ------------------------------------
struct D3 {
    sf::f64 x, y, z;
};
CR_HDI D3 widen(V3 v) { return {sf::from_float(v.x), sf::from_float(v.y), sf::from_float(v.z)}; }
CR_HDI sf::f64 dat(const D3& v, int i) { return i == 0 ? v.x : (i == 1 ? v.y : v.z); }
CR_HDI D3 dadd(const D3& a, const D3& b) { return {sf::add(a.x, b.x), sf::add(a.y, b.y), sf::add(a.z, b.z)}; }
CR_HDI D3 dsub(const D3& a, const D3& b) { return {sf::sub(a.x, b.x), sf::sub(a.y, b.y), sf::sub(a.z, b.z)}; }
CR_HDI D3 dscale(const D3& a, sf::f64 s) { return {sf::mul(a.x, s), sf::mul(a.y, s), sf::mul(a.z, s)}; }
CR_HDI sf::f64 ddot(const D3& a, const D3& b) {
    return sf::add(sf::add(sf::mul(a.x, b.x), sf::mul(a.y, b.y)), sf::mul(a.z, b.z));
}
CR_HDI V3 narrow(const D3& v) { return {sf::to_float(v.x), sf::to_float(v.y), sf::to_float(v.z)}; }

// unit(): double-precision normalization with a fallback for zero/invalid input
// (reference src/geometry.cpp:31-37).
// ---- float-float: an unevaluated pair hi + lo of floats (about 48 bits) ----
struct FF {
    float hi, lo;
};
CR_HDI FF ff_two_sum(float a, float b) {
    const float s = a + b, bb = s - a;
    return FF{s, (a - (s - bb)) + (b - bb)};
}
CR_HDI FF ff_fast_two_sum(float a, float b) {  // |a| >= |b|
    const float s = a + b;
    return FF{s, b - (s - a)};
}
CR_HDI FF ff_mul(FF a, FF b) {
    const float p = a.hi * b.hi;
    const float e = fmaf(a.hi, b.hi, -p) + (a.hi * b.lo + a.lo * b.hi);
    return ff_fast_two_sum(p, e);
}
CR_HDI FF ff_add(FF a, FF b) {
    const FF s = ff_two_sum(a.hi, b.hi);
    return ff_fast_two_sum(s.hi, s.lo + (a.lo + b.lo));
}
CR_HDI FF ff_neg(FF a) { return FF{-a.hi, -a.lo}; }
// a / b for float pairs, to about 2^-46 relative (the remainder of the first
// quotient digit is exact in the fused multiply-add).
CR_HDI FF ff_div2(FF a, FF b) {
    const float q0 = a.hi / b.hi;
    const float r = (fmaf(-q0, b.hi, a.hi) + a.lo) - q0 * b.lo;
    return ff_fast_two_sum(q0, r / b.hi);
}

CR_HDI V3 dunit(const D3& v, V3 fallback) {
    const sf::f64 sq = ddot(v, v);
    if (!sf::gt(sq, sf::zero(0)) || !sf::is_finite(sq)) return fallback;
    const sf::f64 inv = sf::div(sf::one(), sf::sqrt(sq));
    return {sf::to_float(sf::mul(v.x, inv)), sf::to_float(sf::mul(v.y, inv)), sf::to_float(sf::mul(v.z, inv))};
}

// A binary64 value as a float pair holding its leading 48 significand bits
// (relative error below 2^-47); false outside 2^-60..2^60 (and for inf/NaN).
CR_HDI bool ff_of(sf::f64 a, FF& out) {
    const float sgn = sf::sign_of(a) ? -1.0f : 1.0f;
    if (sf::is_zero(a)) {
        out = FF{sgn * 0.0f, 0.0f};
        return true;
    }
    const int32_t e = sf::exp_of(a) - 1023;
    if (e < -60 || e > 60) return false;
    const uint64_t m = sf::frac_of(a) | (uint64_t(1) << 52);
    out.hi = sgn * ldexpf(float(uint32_t(m >> 29)), e - 23);
    out.lo = sgn * ldexpf(float(uint32_t((m >> 5) & 0xffffffu)), e - 47);
    return true;
}

// dunit without the binary64 emulation on the common path: the vector's
// components as float pairs, its length and the three quotients to about
// 2^-44 relative in float-float, each rounded once.  The reference's binary64
// chain is within 2^-50 of the exact quotients, so both round alike unless a
// quotient lies within 2^-38 (relative) of a binary32 rounding boundary;
// there, and for components outside 2^-60..2^60, dunit decides.
// The core of the float-float normalizations below: the quotients
// c_k / |c| rounded once to float, where err[k] bounds the absolute error of
// component k against the binary64 vector the reference normalizes (0: the
// pair holds that component to 2^-47).  A unit vector moves by at most
// 2 |delta| / |c| when its vector moves by delta, so each quotient must lie
// farther than 4 sum(err) / |c| plus 2^-38 (relative) from a binary32
// rounding boundary.  False where the binary64 chain must decide.
CR_HDI bool unit_ff(const FF c[3], const float err[3], float out[3]) {
    const FF sq = ff_add(ff_add(ff_mul(c[0], c[0]), ff_mul(c[1], c[1])), ff_mul(c[2], c[2]));
    if (!(sq.hi > 0x1p-110f)) return false;
    const float s0 = sqrtf(sq.hi);
    const float res = fmaf(-s0, s0, sq.hi) + sq.lo;
    const FF nrm = ff_fast_two_sum(s0, res / (2.0f * s0));
    const float slack = 4.0f * ((err[0] + err[1]) + err[2]) / nrm.hi;
    for (int k = 0; k < 3; ++k) {
        if (c[k].hi == 0.0f) {  // a signed zero stays itself, as in the binary64 product
            if (err[k] != 0.0f) return false;
            out[k] = c[k].hi;
            continue;
        }
        const float q0 = c[k].hi / nrm.hi;
        const float p = q0 * nrm.hi, pe = fmaf(q0, nrm.hi, -p);
        const float r = (((c[k].hi - p) - pe) + c[k].lo) - q0 * nrm.lo;
        const float q1 = r / nrm.hi;
        const float f = q0 + q1, af = fabsf(f);
        const float d = fabsf((q0 - f) + q1);
        const float half_ulp = 0.5f * fminf(nextafterf(af, INFINITY) - af, af - nextafterf(af, 0.0f));
        if (!(af >= 0x1p-100f) || !(half_ulp - d >= 0x1p-38f * af + slack)) return false;
        out[k] = f;
    }
    return true;
}

CR_HDI V3 dunit_ff(const D3& v, V3 fallback) {
    FF c[3];
    if (!ff_of(v.x, c[0]) || !ff_of(v.y, c[1]) || !ff_of(v.z, c[2])) return dunit(v, fallback);
    const float exact[3] = {0.0f, 0.0f, 0.0f};
    float out[3];
    if (!unit_ff(c, exact, out)) return dunit(v, fallback);
    return V3{out[0], out[1], out[2]};
}

// unit(n0 * w + n1 * hu + n2 * hv) for float vectors and binary64 weights, the
// reference evaluating the sum in binary64 ((n0 w + n1 hu) + n2 hv, each step
// rounded): the sum in float-float, whose distance from the reference's is
// below 2^-44 of the terms' magnitudes (plus 2^-120 for products near the
// subnormal range), then unit_ff.  eu, ev bound how far the pairs fu, fv may
// lie from the reference's binary64 weights beyond that (0: the pairs hold
// them to 2^-47), and fw = 1 - fu - fv as a pair errs against the reference's
// (1 - hu) - hv by the negated errors of fu and fv plus at most ew: so the
// weights' errors move the sum by at most |n1 - n0| eu + |n2 - n0| ev + |n0| ew
// per component (an edge-length, not a coordinate, times eu and ev).  A component whose three
// inputs are zero is the signed zero of the binary64 sum.  False where the
// caller's binary64 chain must decide.
CR_HDI bool interp_unit_ff(V3 n0, V3 n1, V3 n2, FF fw, FF fu, FF fv, V3& out, float ew = 0.0f, float eu = 0.0f,
                           float ev = 0.0f) {
    FF c[3];
    float err[3];
    for (int k = 0; k < 3; ++k) {
        const float a = at(n0, k), b = at(n1, k), e = at(n2, k);
        if (a == 0.0f && b == 0.0f && e == 0.0f) {
            c[k] = FF{(a * fw.hi + b * fu.hi) + e * fv.hi, 0.0f};
            err[k] = 0.0f;
            continue;
        }
        const FF t0 = ff_mul(FF{a, 0.0f}, fw), t1 = ff_mul(FF{b, 0.0f}, fu), t2 = ff_mul(FF{e, 0.0f}, fv);
        c[k] = ff_add(ff_add(t0, t1), t2);
        err[k] = 0x1p-44f * ((fabsf(t0.hi) + fabsf(t1.hi)) + fabsf(t2.hi)) + 0x1p-120f +
                 1.01f * ((fabsf(b - a) * eu + fabsf(e - a) * ev) + fabsf(a) * ew);
    }
    float o[3];
    if (!unit_ff(c, err, o)) return false;
    out = V3{o[0], o[1], o[2]};
    return true;
}
CR_HDI bool interp_unit_ff(V3 n0, V3 n1, V3 n2, sf::f64 w, sf::f64 hu, sf::f64 hv, V3& out) {
    FF fw, fu, fv;
    if (!ff_of(w, fw) || !ff_of(hu, fu) || !ff_of(hv, fv)) return false;
    return interp_unit_ff(n0, n1, n2, fw, fu, fv, out);
}

// float((a * w + b * hu) + e * hv) per component for float vectors a, b, e and
// binary64 weights, the sum being the reference's binary64 chain (each step
// rounded): the sum in float-float as in interp_unit_ff (within 2^-44 of the
// terms' magnitudes, plus 2^-120, of the chain's value), kept only where that
// distance (plus the weight errors ew, eu, ev, as in interp_unit_ff) leaves the
// float rounding of the chain decided.  A component whose
// three inputs are zero is the signed zero of the binary64 sum.  False where
// the caller's binary64 chain must decide.
CR_HDI bool interp_point_ff(V3 a, V3 b, V3 e, FF fw, FF fu, FF fv, V3& out, float ew = 0.0f, float eu = 0.0f,
                            float ev = 0.0f) {
    float o[3];
    for (int k = 0; k < 3; ++k) {
        const float x = at(a, k), y = at(b, k), z = at(e, k);
        if (x == 0.0f && y == 0.0f && z == 0.0f) {
            o[k] = (x * fw.hi + y * fu.hi) + z * fv.hi;
            continue;
        }
        const FF t0 = ff_mul(FF{x, 0.0f}, fw), t1 = ff_mul(FF{y, 0.0f}, fu), t2 = ff_mul(FF{z, 0.0f}, fv);
        const FF c = ff_add(ff_add(t0, t1), t2);
        const float err = 0x1p-44f * ((fabsf(t0.hi) + fabsf(t1.hi)) + fabsf(t2.hi)) + 0x1p-120f +
                          1.01f * ((fabsf(y - x) * eu + fabsf(z - x) * ev) + fabsf(x) * ew);
        const float af = fabsf(c.hi);
        const float half_ulp = 0.5f * fminf(nextafterf(af, INFINITY) - af, af - nextafterf(af, 0.0f));
        if (!(af >= 0x1p-100f) || !(af < 0x1p100f) || !(half_ulp - fabsf(c.lo) > 1.01f * err)) return false;
        o[k] = c.hi;
    }
    out = V3{o[0], o[1], o[2]};
    return true;
}

// Sign of the binary64 dot product of two float vectors, exactly as the
// reference evaluates preciseDot(DVec3(a), DVec3(b)).  Float products are
// exact in binary64, so a float-only filter decides almost every call; ties
// near zero fall back to the emulated sum.  Returns -1, 0 or +1.
CR_HDI int dot_sign_exact(V3 a, V3 b) {
    const float p0 = a.x * b.x, p1 = a.y * b.y, p2 = a.z * b.z;
    const float s = (p0 + p1) + p2;
    const float mag = fabsf(p0) + fabsf(p1) + fabsf(p2);
    if (finite1(s) && finite1(mag) && fabsf(s) > mag * 4.0e-7f && mag > 1.0e-30f) return s > 0 ? 1 : -1;
    const sf::f64 d = ddot(widen(a), widen(b));
    if (sf::is_nan(d)) return 0;
    if (sf::is_zero(d)) return 0;
    return sf::sign_of(d) ? -1 : 1;
}

// powerHeuristic in binary64 (reference src/math.h:124-128).
CR_HDI float power_heuristic(float a, float b) {
    if (a <= 0) return 0;
    const sf::f64 da = sf::from_float(a), db = sf::from_float(b);
    const sf::f64 aa = sf::mul(da, da), bb = sf::mul(db, db);
    return sf::to_float(sf::div(aa, sf::add(aa, bb)));
}

}  // namespace cr
