// sf64.h — exact IEEE-754 binary64 arithmetic from 64-bit integer operations.
//
// The reference renderer evaluates its geometric predicates in double
// (reference src/geometry.cpp:147-182 intersectTriangle, :186-209 offsetRay,
// :213-234 Camera::generate, :401-459 surfaceAt; src/math.h:124-128
// powerHeuristic; src/render.cpp:390-415 film accumulation).  The device path
// executes no FP64 instructions; wherever a result must equal the reference's
// double result bit for bit it uses this emulation instead.  Every operation
// is correctly rounded (round-to-nearest-even) including subnormal results,
// signed zeros, infinities and NaN, so `sf64 a*b` equals `double a*b` for all
// inputs.  The same source compiles for the host, where tests/test_sf64*.py
// checks it against native double on random and edge-case operands.
#pragma once
#include <math.h>
#include <stdint.h>

#ifndef CR_HD
#if defined(__HIPCC__) || defined(__HIP__)
#define CR_HD __host__ __device__
#define CR_HDI __host__ __device__ inline
#define CR_NOINLINE __noinline__
#else
#define CR_HD
#define CR_HDI inline
#define CR_NOINLINE __attribute__((noinline))
#endif
#endif

// Large routines called from several sites.  Inlined: keeping one out-of-line
// copy of each (smaller kernels) measured slower on gfx950 (call overhead
// outweighs the instruction-cache gain for the shading kernels).
#ifndef CR_OUTLINE
#define CR_OUTLINE CR_HD inline
#endif

namespace sf {

struct f64 {
    uint64_t b;
};

CR_HDI f64 bits(uint64_t b) { return f64{b}; }
CR_HDI uint32_t sign_of(f64 a) { return uint32_t(a.b >> 63); }
CR_HDI int32_t exp_of(f64 a) { return int32_t((a.b >> 52) & 0x7ff); }
CR_HDI uint64_t frac_of(f64 a) { return a.b & 0x000fffffffffffffull; }
CR_HDI bool is_nan(f64 a) { return exp_of(a) == 0x7ff && frac_of(a) != 0; }
CR_HDI bool is_inf(f64 a) { return (a.b & 0x7fffffffffffffffull) == 0x7ff0000000000000ull; }
CR_HDI bool is_finite(f64 a) { return exp_of(a) != 0x7ff; }
CR_HDI bool is_zero(f64 a) { return (a.b & 0x7fffffffffffffffull) == 0; }
CR_HDI f64 neg(f64 a) { return f64{a.b ^ 0x8000000000000000ull}; }
CR_HDI f64 fabs_(f64 a) { return f64{a.b & 0x7fffffffffffffffull}; }
CR_HDI f64 qnan() { return f64{0x7ff8000000000000ull}; }
CR_HDI f64 inf(uint32_t s) { return f64{(uint64_t(s) << 63) | 0x7ff0000000000000ull}; }
CR_HDI f64 zero(uint32_t s) { return f64{uint64_t(s) << 63}; }

CR_HDI int clz64(uint64_t x) { return x ? __builtin_clzll(x) : 64; }

CR_HDI uint64_t shift_right_jam(uint64_t x, int n) {
    if (n <= 0) return x;
    if (n >= 64) return x != 0;
    return (x >> n) | uint64_t((x << (64 - n)) != 0);
}

CR_HDI uint64_t mul_hi(uint64_t a, uint64_t b) { return uint64_t((unsigned __int128)a * b >> 64); }

// Round and pack. `sig` holds the significand with its leading one at bit 62
// (bits 9..0 are guard/sticky bits); the value is sig * 2^(e - 62).
CR_HDI f64 round_pack(uint32_t s, int32_t e, uint64_t sig) {
    int32_t be = e + 1023;
    if (be >= 0x7ff) return inf(s);
    bool sub = false;
    if (be <= 0) {
        sig = shift_right_jam(sig, 1 - be);
        be = 0;
        sub = true;
    }
    const uint32_t rb = uint32_t(sig & 0x3ff);
    sig = (sig + 0x200) >> 10;
    if (rb == 0x200) sig &= ~uint64_t(1);
    if (sub) return f64{(uint64_t(s) << 63) + sig};
    return f64{(uint64_t(s) << 63) + (uint64_t(be - 1) << 52) + sig};
}

// Normalize a nonzero finite operand: sig gets its leading one at bit 52.
CR_HDI void unpack(f64 a, int32_t& e, uint64_t& sig) {
    const int32_t ex = exp_of(a);
    const uint64_t fr = frac_of(a);
    if (ex == 0) {
        const int sh = clz64(fr) - 11;
        sig = fr << sh;
        e = -1022 - sh;
    } else {
        sig = fr | 0x0010000000000000ull;
        e = ex - 1023;
    }
}

CR_HDI f64 add_mags(f64 a, f64 b, uint32_t s) {
    // |a| + |b| with both finite and nonzero.
    int32_t ea, eb;
    uint64_t sa, sb;
    unpack(a, ea, sa);
    unpack(b, eb, sb);
    if (ea < eb) {
        int32_t t = ea; ea = eb; eb = t;
        uint64_t u = sa; sa = sb; sb = u;
    }
    sa <<= 10;
    sb = shift_right_jam(sb << 10, ea - eb);
    uint64_t sum = sa + sb;
    if (sum >> 63) {
        sum = shift_right_jam(sum, 1);
        ++ea;
    }
    return round_pack(s, ea, sum);
}

CR_HDI f64 sub_mags(f64 a, f64 b, uint32_t s) {
    // |a| - |b| carrying sign s for a positive result; both finite nonzero.
    int32_t ea, eb;
    uint64_t sa, sb;
    unpack(a, ea, sa);
    unpack(b, eb, sb);
    if (ea < eb || (ea == eb && sa < sb)) {
        int32_t t = ea; ea = eb; eb = t;
        uint64_t u = sa; sa = sb; sb = u;
        s ^= 1;
    }
    if (ea == eb && sa == sb) return zero(0);
    sa <<= 10;
    sb = shift_right_jam(sb << 10, ea - eb);
    uint64_t d = sa - sb;
    const int sh = clz64(d) - 1;
    d <<= sh;
    return round_pack(s, ea - sh, d);
}

// General case of add (special values, subnormal operands or result).
CR_HDI f64 add_general(f64 a, f64 b) {
    if (is_nan(a) || is_nan(b)) return qnan();
    const uint32_t sa = sign_of(a), sb = sign_of(b);
    if (is_inf(a)) {
        if (is_inf(b) && sa != sb) return qnan();
        return a;
    }
    if (is_inf(b)) return b;
    if (is_zero(a)) {
        if (is_zero(b)) return zero(sa & sb);
        return b;
    }
    if (is_zero(b)) return a;
    return sa == sb ? add_mags(a, b, sa) : sub_mags(a, b, sa);
}

// Round-to-nearest-even of a significand with its leading one at bit 62 and
// 10 guard bits (bit 0 sticky), packed with biased exponent e >= 1; a carry out
// of the significand moves into the exponent (0x7ff: infinity).
CR_HDI f64 pack_normal(uint64_t sign, int32_t e, uint64_t sig) {
    const uint64_t rb = sig & 0x3ff;
    uint64_t m = sig >> 10;
    m += (rb > 0x200 || (rb == 0x200 && (m & 1))) ? 1 : 0;
    return f64{sign | ((uint64_t(e - 1) << 52) + m)};
}

// Common case of both additions and subtractions: normal operands whose sum
// is normal (no special value, no subnormal operand or result); everything else
// goes to add_general.  Same rounding as the general path.
CR_HD CR_NOINLINE inline f64 add(f64 a, f64 b) {
    const uint32_t ea = uint32_t(a.b >> 52) & 0x7ffu, eb = uint32_t(b.b >> 52) & 0x7ffu;
    if (ea - 1u < 0x7fdu && eb - 1u < 0x7fdu) {  // exponents in [1, 0x7fd]
        const uint64_t kMag = 0x7fffffffffffffffull, kFrac = 0x000fffffffffffffull, kHid = 0x0010000000000000ull;
        const bool swap = (b.b & kMag) > (a.b & kMag);
        const uint64_t x = swap ? b.b : a.b, y = swap ? a.b : b.b;
        const int32_t ex = int32_t((x >> 52) & 0x7ff), ey = int32_t((y >> 52) & 0x7ff);
        const uint64_t sx = ((x & kFrac) | kHid) << 10;
        const uint64_t sy = shift_right_jam(((y & kFrac) | kHid) << 10, ex - ey);
        const uint64_t sign = x & ~kMag;
        if (((a.b ^ b.b) >> 63) == 0) {
            uint64_t s_ = sx + sy;
            int32_t e = ex;
            if (s_ >> 63) {
                s_ = (s_ >> 1) | (s_ & 1);
                ++e;
            }
            return pack_normal(sign, e, s_);
        }
        uint64_t d = sx - sy;
        if (d == 0) return zero(0);
        const int sh = clz64(d) - 1;
        if (ex - sh >= 1) return pack_normal(sign, ex - sh, d << sh);
    }
    return add_general(a, b);
}

CR_HDI f64 sub(f64 a, f64 b) { return add(a, is_nan(b) ? b : neg(b)); }

// General case of mul (special values, subnormal operands or result).
CR_HDI f64 mul_general(f64 a, f64 b) {
    if (is_nan(a) || is_nan(b)) return qnan();
    const uint32_t s = sign_of(a) ^ sign_of(b);
    if (is_inf(a) || is_inf(b)) {
        if (is_zero(a) || is_zero(b)) return qnan();
        return inf(s);
    }
    if (is_zero(a) || is_zero(b)) return zero(s);
    int32_t ea, eb;
    uint64_t ma, mb;
    unpack(a, ea, ma);
    unpack(b, eb, mb);
    ma <<= 10;
    mb <<= 11;
    const uint64_t hi = mul_hi(ma, mb), lo = ma * mb;
    uint64_t sig = hi | uint64_t(lo != 0);
    int32_t e = ea + eb + 1;
    if (sig < 0x4000000000000000ull) {
        sig <<= 1;
        --e;
    }
    return round_pack(s, e, sig);
}

// Common case: normal operands with a normal product (exponent range checked
// before the multiply); everything else goes to mul_general.
CR_HD CR_NOINLINE inline f64 mul(f64 a, f64 b) {
    const uint32_t ea = uint32_t(a.b >> 52) & 0x7ffu, eb = uint32_t(b.b >> 52) & 0x7ffu;
    const int32_t e0 = int32_t(ea + eb) - 1023;  // biased exponent when the significand product is < 2
    if (ea - 1u < 0x7feu && eb - 1u < 0x7feu && e0 >= 1 && e0 <= 0x7fd) {
        const uint64_t kFrac = 0x000fffffffffffffull, kHid = 0x0010000000000000ull;
        const uint64_t ma = ((a.b & kFrac) | kHid) << 10, mb = ((b.b & kFrac) | kHid) << 11;
        uint64_t sig = mul_hi(ma, mb) | uint64_t(ma * mb != 0);  // leading one at bit 62 or 61
        int32_t e = e0 + 1;
        if (sig < 0x4000000000000000ull) {
            sig <<= 1;
            --e;
        }
        return pack_normal((a.b ^ b.b) & 0x8000000000000000ull, e, sig);
    }
    return mul_general(a, b);
}

// Integer -> float conversions used for quotient-digit estimates.
CR_HDI float u64_to_f(uint64_t x) { return float(x); }
CR_HDI float i64_to_f(int64_t x) { return float(x); }

// Correctly rounded quotient.  With A in [B, 2B) and B in [2^52, 2^53), the
// 63-bit quotient Q = floor(A*2^62/B) is produced as three signed digits
// (23, 23 and 16 bits) from a float reciprocal estimate; each partial
// remainder R_k = R_{k-1}*2^s - Q_k*B is computed exactly modulo 2^64 (its
// true value is below 2^58 in magnitude, so wrapping arithmetic is exact),
// and a final loop brings the remainder into [0, B).  Digit errors are <= 4,
// <= 17 and <= 2 units, so the loop runs at most a few times.
// Quotient significand for ma in [mb, 2*mb), mb in [2^52, 2^53): Q with its
// leading one at bit 62 and the remainder's sticky bit in bit 0.
CR_HDI uint64_t quotient_sig(uint64_t ma, uint64_t mb) {
    const float rb = 1.0f / u64_to_f(mb);  // ~2^-53, relative error <= 2^-23
    // digit 1 (weight 2^39): ~A*2^23/B
    const int64_t q1 = int64_t(u64_to_f(ma) * (rb * 8388608.0f));
    int64_t r = int64_t((ma << 23) - uint64_t(q1) * mb);
    // digit 2 (weight 2^16)
    const int64_t q2 = int64_t(i64_to_f(r) * (rb * 8388608.0f));
    r = int64_t((uint64_t(r) << 23) - uint64_t(q2) * mb);
    // digit 3 (weight 1)
    int64_t q3 = int64_t(i64_to_f(r) * (rb * 65536.0f));
    r = int64_t((uint64_t(r) << 16) - uint64_t(q3) * mb);
    while (r < 0) {
        --q3;
        r += int64_t(mb);
    }
    while (r >= int64_t(mb)) {
        ++q3;
        r -= int64_t(mb);
    }
    const uint64_t q = (uint64_t(q1) << 39) + (uint64_t(q2) << 16) + uint64_t(q3);
    return q | uint64_t(r != 0);
}

// One out-of-line copy (fewer instructions to fetch in the shading kernel;
// measured slightly faster than inlining every division).
CR_HD CR_NOINLINE inline f64 div(f64 a, f64 b) {
    const uint32_t xa = uint32_t(a.b >> 52) & 0x7ffu, xb = uint32_t(b.b >> 52) & 0x7ffu;
    const uint32_t s = sign_of(a) ^ sign_of(b);
    int32_t e;
    uint64_t ma, mb;
    if (xa - 1u < 0x7feu && xb - 1u < 0x7feu) {  // normal operands: no special case, no unpack shift
        ma = frac_of(a) | 0x0010000000000000ull;
        mb = frac_of(b) | 0x0010000000000000ull;
        e = int32_t(xa) - int32_t(xb);
    } else {
        if (is_nan(a) || is_nan(b)) return qnan();
        if (is_inf(a)) return is_inf(b) ? qnan() : inf(s);
        if (is_inf(b)) return zero(s);
        if (is_zero(b)) return is_zero(a) ? qnan() : inf(s);
        if (is_zero(a)) return zero(s);
        int32_t ea, eb;
        unpack(a, ea, ma);
        unpack(b, eb, mb);
        e = ea - eb;
    }
    if (ma < mb) {
        ma <<= 1;
        --e;
    }
    const uint64_t q = quotient_sig(ma, mb);
    const int32_t be = e + 1023;
    if (be >= 1 && be <= 0x7fe) return pack_normal(uint64_t(s) << 63, be, q);
    return round_pack(s, e, q);
}

// 128-bit helpers for the square root.
struct U128 {
    uint64_t hi, lo;
};
CR_HDI U128 sq128(uint64_t x) { return U128{mul_hi(x, x), x * x}; }
// Signed difference N - X of two 128-bit values whose difference fits in 127 bits.
CR_HDI void sub128(U128 n, U128 x, int64_t& hi, uint64_t& lo) {
    lo = n.lo - x.lo;
    hi = int64_t(n.hi - x.hi - (n.lo < x.lo ? 1 : 0));
}
CR_HDI float i128_to_f(int64_t hi, uint64_t lo) { return float(hi) * 18446744073709551616.0f + float(lo); }

// Correctly rounded square root.  For m in [2^52, 2^54) the 55-bit integer root
// R = floor(sqrt(m*2^56)) is reached from a float sqrt estimate (error < 2^32)
// by two Newton corrections R += (N - R^2)/(2R) evaluated with float digit
// estimates on the exact 128-bit remainder, then a final +-1 adjustment.
CR_OUTLINE f64 sqrt(f64 a) {
    if (is_nan(a)) return qnan();
    if (is_zero(a)) return a;
    if (sign_of(a)) return qnan();
    if (is_inf(a)) return a;
    int32_t e;
    uint64_t m;
    unpack(a, e, m);
    if (e & 1) {
        m <<= 1;
        --e;
    }
    const U128 n{m >> 8, m << 56};  // N = m * 2^56
    uint64_t root = uint64_t(sqrtf(u64_to_f(m))) << 28;
    for (int it = 0; it < 2; ++it) {
        int64_t dh;
        uint64_t dl;
        sub128(n, sq128(root), dh, dl);
        const float corr = i128_to_f(dh, dl) / (2.0f * u64_to_f(root));
        root = uint64_t(int64_t(root) + int64_t(corr));
    }
    for (;;) {  // root^2 <= N < (root+1)^2
        int64_t dh;
        uint64_t dl;
        sub128(n, sq128(root), dh, dl);
        if (dh < 0) {
            --root;
            continue;
        }
        const uint64_t r1 = root + 1;
        sub128(n, sq128(r1), dh, dl);
        if (dh >= 0) {  // (root+1)^2 <= N
            ++root;
            continue;
        }
        break;
    }
    int64_t dh;
    uint64_t dl;
    sub128(n, sq128(root), dh, dl);
    const uint64_t sig = (root << 8) | uint64_t(dh != 0 || dl != 0);
    return round_pack(0, e / 2, sig);
}

// ---- comparisons (IEEE semantics: NaN compares false, -0 == +0) ----------
CR_HDI int64_t order_key(f64 a) {
    // Monotone integer key for non-NaN values.
    return sign_of(a) ? -int64_t(a.b & 0x7fffffffffffffffull) : int64_t(a.b);
}
CR_HDI bool lt(f64 a, f64 b) {
    if (is_nan(a) || is_nan(b)) return false;
    if (is_zero(a) && is_zero(b)) return false;
    return order_key(a) < order_key(b);
}
CR_HDI bool le(f64 a, f64 b) {
    if (is_nan(a) || is_nan(b)) return false;
    if (is_zero(a) && is_zero(b)) return true;
    return order_key(a) <= order_key(b);
}
CR_HDI bool eq(f64 a, f64 b) {
    if (is_nan(a) || is_nan(b)) return false;
    if (is_zero(a) && is_zero(b)) return true;
    return a.b == b.b;
}
CR_HDI bool gt(f64 a, f64 b) { return lt(b, a); }
CR_HDI bool ge(f64 a, f64 b) { return le(b, a); }
// std::max / std::min semantics: max(a,b) = (a < b) ? b : a.
CR_HDI f64 max_(f64 a, f64 b) { return lt(a, b) ? b : a; }
CR_HDI f64 min_(f64 a, f64 b) { return lt(b, a) ? b : a; }

// ---- conversions -----------------------------------------------------------
CR_HDI uint32_t fbits(float x) {
    union { float f; uint32_t u; } c;
    c.f = x;
    return c.u;
}
CR_HDI float bitsf(uint32_t u) {
    union { float f; uint32_t u; } c;
    c.u = u;
    return c.f;
}

// Exact widening float -> double.
CR_HDI f64 from_float(float x) {
    const uint32_t u = fbits(x);
    const uint32_t s = u >> 31;
    const int32_t ex = int32_t((u >> 23) & 0xff);
    uint32_t fr = u & 0x7fffff;
    if (ex == 0xff) {
        if (fr) return qnan();
        return inf(s);
    }
    if (ex == 0) {
        if (fr == 0) return zero(s);
        // float subnormal: normalize
        int32_t e = -126;
        while (!(fr & 0x800000)) {
            fr <<= 1;
            --e;
        }
        fr &= 0x7fffff;
        return f64{(uint64_t(s) << 63) | (uint64_t(e + 1023) << 52) | (uint64_t(fr) << 29)};
    }
    return f64{(uint64_t(s) << 63) | (uint64_t(ex - 127 + 1023) << 52) | (uint64_t(fr) << 29)};
}

// double(x) - double(y) for binary32 x, y.  When both are normal and their
// exponents differ by at most 28, the difference of the two 24-bit significands
// aligned to the smaller exponent has at most 53 bits, so it is formed exactly
// in a 64-bit integer and packed without rounding; every other case (zeros,
// subnormals, infinities, NaN, wider exponent gaps) takes the general path.
CR_HDI f64 sub_floats(float x, float y) {
    const uint32_t ux = fbits(x), uy = fbits(y);
    const int32_t ex = int32_t((ux >> 23) & 0xffu), ey = int32_t((uy >> 23) & 0xffu);
    const int32_t gap = ex - ey;
    if (uint32_t(ex - 1) < 254u && uint32_t(ey - 1) < 254u && gap <= 28 && gap >= -28) {
        const int32_t emin = ex < ey ? ex : ey;
        int64_t X = int64_t((ux & 0x7fffffu) | 0x800000u) << (ex - emin);
        int64_t Y = int64_t((uy & 0x7fffffu) | 0x800000u) << (ey - emin);
        if (ux >> 31) X = -X;
        if (uy >> 31) Y = -Y;
        const int64_t D = X - Y;  // |D| < 2^53
        if (D == 0) return zero(0);
        const uint64_t mag = D < 0 ? uint64_t(-D) : uint64_t(D);
        const int32_t lead = 63 - clz64(mag);  // <= 52
        // D * 2^(emin - 150); biased binary64 exponent emin - 150 + lead + 1023
        return f64{(uint64_t(D < 0) << 63) | (uint64_t(emin + 873 + lead) << 52) |
                   ((mag << (52 - lead)) & 0x000fffffffffffffull)};
    }
    return sub(from_float(x), from_float(y));
}

// Correctly rounded narrowing double -> float (round-to-nearest-even).
CR_HD CR_NOINLINE inline float to_float(f64 a) {
    const uint32_t xa = uint32_t(a.b >> 52) & 0x7ffu;
    if (xa - 897u < 254u) {  // binary32 exponent xa - 896 in [1, 254]: normal result
        const uint64_t m = frac_of(a) | 0x0010000000000000ull;
        const uint64_t rb = m & ((uint64_t(1) << 29) - 1), half = uint64_t(1) << 28;
        uint32_t r = uint32_t(m >> 29);
        r += (rb > half || (rb == half && (r & 1u))) ? 1u : 0u;
        return bitsf((uint32_t(a.b >> 63) << 31) | ((uint32_t(xa - 897u) << 23) + r));
    }
    const uint32_t s = sign_of(a);
    if (is_nan(a)) return bitsf(0x7fc00000u);
    if (is_inf(a)) return bitsf((s << 31) | 0x7f800000u);
    if (is_zero(a)) return bitsf(s << 31);
    int32_t e;
    uint64_t m;
    unpack(a, e, m);
    uint64_t sig = m << 10;  // leading one at bit 62
    int32_t be = e + 127;
    if (be >= 0xff) return bitsf((s << 31) | 0x7f800000u);
    bool sub = false;
    if (be <= 0) {
        sig = shift_right_jam(sig, 1 - be);
        be = 0;
        sub = true;
    }
    const uint64_t half = uint64_t(1) << 38;
    const uint64_t rb = sig & ((uint64_t(1) << 39) - 1);
    uint64_t r = (sig + half) >> 39;
    if (rb == half) r &= ~uint64_t(1);
    if (sub) return bitsf((s << 31) + uint32_t(r));
    const uint64_t packed = (uint64_t(s) << 31) + (uint64_t(be - 1) << 23) + r;
    if (((packed >> 23) & 0xff) == 0xff) return bitsf((s << 31) | 0x7f800000u);
    return bitsf(uint32_t(packed));
}

CR_HDI f64 from_u64(uint64_t v) {
    if (v == 0) return zero(0);
    const int lz = clz64(v);
    uint64_t sig;
    if (lz >= 1) sig = v << (lz - 1);  // leading one at bit 62
    else sig = shift_right_jam(v, 1);
    return round_pack(0, 63 - lz, sig);
}
CR_HDI f64 from_i32(int32_t v) {
    if (v < 0) return neg(from_u64(uint64_t(-int64_t(v))));
    return from_u64(uint64_t(v));
}

// Convenience constants.
CR_HDI f64 one() { return f64{0x3ff0000000000000ull}; }
CR_HDI f64 two() { return f64{0x4000000000000000ull}; }

// floor() of a binary64 value (exact; signed zeros and integers pass through).
CR_HDI f64 floor_(f64 a) {
    if (!is_finite(a) || is_zero(a)) return a;
    const int32_t ex = exp_of(a) - 1023;
    if (ex >= 52) return a;
    if (ex < 0) return sign_of(a) ? f64{0xbff0000000000000ull} : zero(0);  // (-1, 0) -> -1, (0, 1) -> +0
    const uint64_t frac = (uint64_t(1) << (52 - ex)) - 1;
    if ((a.b & frac) == 0) return a;
    const f64 t{a.b & ~frac};  // toward zero
    return sign_of(a) ? sub(t, one()) : t;
}

// fmod(a, 2.0) (exact; the result has a's sign and |result| < 2).
CR_HDI f64 fmod2(f64 a) {
    if (!is_finite(a)) return qnan();
    if (is_zero(a)) return a;
    const int32_t ex = exp_of(a) - 1023;
    if (ex < 1) return a;                     // |a| < 2
    if (ex >= 53) return zero(sign_of(a));    // an even integer
    const f64 half{a.b - (uint64_t(1) << 52)};  // a / 2 (exact: a is normal with ex >= 1)
    const int32_t hx = ex - 1;
    const uint64_t frac = hx >= 52 ? 0 : (uint64_t(1) << (52 - hx)) - 1;
    const f64 q{half.b & ~frac};                // trunc(a / 2)
    const f64 r = sub(a, f64{q.b + (uint64_t(1) << 52)});  // a - 2*trunc(a/2), exact
    return is_zero(r) ? zero(sign_of(a)) : r;
}

// Integral binary64 value with |a| < 2^62 to int64.
CR_HDI int64_t to_i64(f64 a) {
    if (is_zero(a)) return 0;
    int32_t e;
    uint64_t m;
    unpack(a, e, m);
    const uint64_t mag = e >= 52 ? m << (e - 52) : m >> (52 - e);
    return sign_of(a) ? -int64_t(mag) : int64_t(mag);
}


// e^x of a binary32 argument, rounded once to binary32.  Binary64 reduction
// x = n*ln2 + r with ln2 split into a 32-bit head (n*head and x - n*head are
// exact for |n| < 2^8) and a binary64 tail, Taylor series of e^r to r^13 for
// |r| <= ln2/2 (truncation below 2^-56), scaling by 2^n, and one rounding to
// binary32.  The binary64 value is within about 2^-51 relative of e^x, so the
// result is e^x correctly rounded except where e^x lies that close to a
// binary32 rounding boundary.
CR_HD CR_NOINLINE inline float expf_cr(float xf) {
    if (!(xf == xf)) return xf;
    if (xf < -104.0f) return 0.0f;       // e^x < 2^-150 rounds to +0
    if (xf > 89.0f) return bitsf(0x7f800000u);
    const f64 x = from_float(xf);
    const f64 t = mul(x, f64{0x3ff71547652b82feull});                       // x / ln2
    const int32_t n = int32_t(to_i64(floor_(add(t, f64{0x3fe0000000000000ull}))));
    const f64 fn = from_i32(n);
    const f64 r = sub(sub(x, mul(fn, f64{0x3fe62e42fee00000ull})), mul(fn, f64{0x3dea39ef35793c76ull}));
    // 1/k! for k = 13 .. 2, then 1, 1.
    constexpr uint64_t kInvFact[12] = {0x3de6124613a86d09ull, 0x3e21eed8eff8d898ull, 0x3e5ae64567f544e4ull,
                                       0x3e927e4fb7789f5cull, 0x3ec71de3a556c734ull, 0x3efa01a01a01a01aull,
                                       0x3f2a01a01a01a01aull, 0x3f56c16c16c16c17ull, 0x3f81111111111111ull,
                                       0x3fa5555555555555ull, 0x3fc5555555555555ull, 0x3fe0000000000000ull};
    f64 p{kInvFact[0]};
    for (int k = 1; k < 12; ++k) p = add(mul(p, r), f64{kInvFact[k]});
    p = add(mul(p, r), one());
    p = add(mul(p, r), one());
    return to_float(mul(p, f64{uint64_t(1023 + n) << 52}));
}

}  // namespace sf
