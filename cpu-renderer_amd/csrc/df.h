// df.h — double-float (unevaluated float pair) arithmetic for filters.
//
// A DF value is hi + lo with |lo| <= ulp(hi)/2.  With a fused multiply-add
// the error-free transforms below are exact, and the composite operations
// have relative errors of a few u^2 (u = 2^-24) as long as no intermediate
// underflows or overflows (Joldes, Muller, Popescu, "Tight and rigorous error
// bounds for basic building blocks of double-word arithmetic", ACM TOMS 2017:
// accurate addition 3u^2, multiplication 5u^2, division 10u^2, square root
// 4u^2).  Callers use these values only inside filters that carry an explicit
// error bound (kDfRel is a generous per-operation relative bound) and fall
// back to the exact binary64 emulation whenever the bound cannot decide a
// result, so a DF result never replaces a binary64 one unless both round to
// the same float.
#pragma once
#include <math.h>

#ifndef CR_HD
#if defined(__HIPCC__) || defined(__HIP__)
#define CR_HD __host__ __device__
#define CR_HDI __host__ __device__ inline
#else
#define CR_HD
#define CR_HDI inline
#endif
#endif

namespace df {

// Per-operation relative error bound used by the filters (true bounds are
// <= 10u^2 ~ 2^-44.7; this leaves a wide margin).
constexpr float kDfRel = 1.0f / 4398046511104.0f;  // 2^-42

struct DF {
    float hi, lo;
};

CR_HDI DF two_sum(float a, float b) {
    const float s = a + b;
    const float bp = s - a;
    const float e = (a - (s - bp)) + (b - bp);
    return DF{s, e};
}

CR_HDI DF fast_two_sum(float a, float b) {  // requires |a| >= |b| (or a == 0)
    const float s = a + b;
    return DF{s, b - (s - a)};
}

CR_HDI DF two_prod(float a, float b) {
    const float p = a * b;
    return DF{p, fmaf(a, b, -p)};
}

CR_HDI DF from(float a) { return DF{a, 0.0f}; }
CR_HDI DF neg(DF a) { return DF{-a.hi, -a.lo}; }

CR_HDI DF add(DF x, DF y) {  // accurate double-word addition
    DF s = two_sum(x.hi, y.hi);
    const DF t = two_sum(x.lo, y.lo);
    s = fast_two_sum(s.hi, s.lo + t.hi);
    return fast_two_sum(s.hi, s.lo + t.lo);
}

CR_HDI DF sub(DF x, DF y) { return add(x, neg(y)); }

CR_HDI DF mul(DF x, DF y) {  // double-word times double-word with FMA
    const DF c = two_prod(x.hi, y.hi);
    float tl = x.lo * y.lo;
    tl = fmaf(x.hi, y.lo, tl);
    tl = fmaf(x.lo, y.hi, tl);
    return fast_two_sum(c.hi, c.lo + tl);
}

CR_HDI DF mulf(DF x, float y) {  // double-word times float
    const DF c = two_prod(x.hi, y);
    return fast_two_sum(c.hi, fmaf(x.lo, y, c.lo));
}

CR_HDI DF div(DF x, DF y) {  // double-word division (one Newton correction)
    const float th = x.hi / y.hi;
    const DF r = sub(x, mulf(y, th));
    const float tl = r.hi / y.hi;
    return fast_two_sum(th, tl);
}

CR_HDI DF sqrt_(DF x) {  // x > 0
    const float s = sqrtf(x.hi);
    const DF r = sub(x, two_prod(s, s));
    return fast_two_sum(s, r.hi / (2.0f * s));
}

CR_HDI float absf(float a) { return fabsf(a); }
CR_HDI float mag(DF a) { return fabsf(a.hi) + fabsf(a.lo); }

}  // namespace df
