// Host check of cpu-renderer_amd/csrc/sf64.h against native binary64.
// Random operands (uniform bit patterns, float-derived values, near-cancelling
// pairs, tiny/huge exponents) plus fixed edge cases; every result must match
// the hardware double bit for bit (NaNs compared as "both NaN").
#include "../../cpu-renderer_amd/csrc/sf64.h"
#include <cmath>
#include <cstdio>
#include <cstring>
#include <random>

static uint64_t db(double d) { uint64_t u; std::memcpy(&u, &d, 8); return u; }
static double bd(uint64_t u) { double d; std::memcpy(&d, &u, 8); return d; }
static bool same(double ref, sf::f64 got) {
    if (std::isnan(ref)) return sf::is_nan(got);
    return db(ref) == got.b;
}
static int failures = 0;
static void report(const char* op, double a, double b, double ref, sf::f64 got) {
    if (failures++ < 20)
        std::printf("FAIL %s a=%a b=%a ref=%a got=%a\n", op, a, b, ref, bd(got.b));
}

int main(int argc, char** argv) {
    const long n = argc > 1 ? std::atol(argv[1]) : 2000000;
    std::mt19937_64 g(12345);
    auto pick = [&](int mode) -> double {
        switch (mode) {
            case 0: return bd(g());                                              // any bit pattern
            case 1: { float f; uint32_t u = uint32_t(g()); std::memcpy(&f, &u, 4); return f; }  // any float
            case 2: return std::ldexp(double(int64_t(g() >> 11)) , int(g() % 200) - 100 - 53);
            case 3: return (g() & 1 ? -1.0 : 1.0) * std::ldexp(1.0 + double(g() >> 12) * 0x1p-52, int(g() % 2100) - 1075);
            default: { float f = float(int(g() % 2001) - 1000) * 0.001f; return f; }
        }
    };
    const double special[] = {0.0, -0.0, 1.0, -1.0, INFINITY, -INFINITY, NAN, 0x1p-1074, -0x1p-1074, 0x1p-1022,
                              0x1.fffffffffffffp1023, 0x1p-1023, 3.0, 0.1, 1e-300, 1e300, 0x1.0000000000001p0};
    for (double a : special) for (double b : special) {
        sf::f64 x{db(a)}, y{db(b)};
        if (!same(a + b, sf::add(x, y))) report("add", a, b, a + b, sf::add(x, y));
        if (!same(a - b, sf::sub(x, y))) report("sub", a, b, a - b, sf::sub(x, y));
        if (!same(a * b, sf::mul(x, y))) report("mul", a, b, a * b, sf::mul(x, y));
        if (!same(a / b, sf::div(x, y))) report("div", a, b, a / b, sf::div(x, y));
        if (sf::lt(x, y) != (a < b) || sf::le(x, y) != (a <= b) || sf::eq(x, y) != (a == b)) report("cmp", a, b, 0, x);
    }
    for (long i = 0; i < n; ++i) {
        const int ma = int(g() % 5), mb = int(g() % 5);
        double a = pick(ma), b = pick(mb);
        if (i % 7 == 0) b = -a * (1.0 + std::ldexp(double(int(g() % 64) - 32), -52));  // cancellation
        sf::f64 x{db(a)}, y{db(b)};
        if (!same(a + b, sf::add(x, y))) report("add", a, b, a + b, sf::add(x, y));
        if (!same(a - b, sf::sub(x, y))) report("sub", a, b, a - b, sf::sub(x, y));
        if (!same(a * b, sf::mul(x, y))) report("mul", a, b, a * b, sf::mul(x, y));
        if (!same(a / b, sf::div(x, y))) report("div", a, b, a / b, sf::div(x, y));
        if (!same(std::sqrt(std::fabs(a)), sf::sqrt(sf::fabs_(x)))) report("sqrt", a, 0, std::sqrt(std::fabs(a)), sf::sqrt(sf::fabs_(x)));
        if (sf::lt(x, y) != (a < b) || sf::le(x, y) != (a <= b) || sf::eq(x, y) != (a == b)) report("cmp", a, b, 0, x);
        const float fa = float(a);
        const float got = sf::to_float(x);
        if (!(std::isnan(fa) ? std::isnan(got) : std::memcmp(&fa, &got, 4) == 0)) report("to_float", a, 0, fa, sf::from_float(got));
        float f; uint32_t u = uint32_t(g()); std::memcpy(&f, &u, 4);
        if (!same(double(f), sf::from_float(f))) report("from_float", f, 0, f, sf::from_float(f));
        const uint64_t iv = g() >> (g() % 64);
        if (!same(double(iv), sf::from_u64(iv))) report("from_u64", double(iv), 0, double(iv), sf::from_u64(iv));
        const int32_t ii = int32_t(g());
        if (!same(double(ii), sf::from_i32(ii))) report("from_i32", ii, 0, ii, sf::from_i32(ii));
    }
    // Division / square-root edge families: exact quotients, near-equal
    // operands, perfect squares and their neighbours.
    for (long i = 0; i < n / 4; ++i) {
        const double b = std::ldexp(1.0 + double(g() >> 12) * 0x1p-52, int(g() % 200) - 100);
        const double q = std::ldexp(double(g() >> (11 + g() % 40)), -int(g() % 30));
        const double a = q * b;  // often exactly representable
        const double near = b * (1.0 + std::ldexp(double(int(g() % 9) - 4), -52));
        const double pairs[3][2] = {{a, b}, {near, b}, {b, near}};
        for (const auto& pq : pairs) {
            sf::f64 x{db(pq[0])}, y{db(pq[1])};
            if (!same(pq[0] / pq[1], sf::div(x, y))) report("div-edge", pq[0], pq[1], pq[0] / pq[1], sf::div(x, y));
        }
        const double r = std::ldexp(double((g() >> 38) | 1), int(g() % 100) - 50);
        for (double v : {r * r, std::nextafter(r * r, 0.0), std::nextafter(r * r, INFINITY)}) {
            sf::f64 x{db(v)};
            if (!same(std::sqrt(v), sf::sqrt(x))) report("sqrt-edge", v, 0, std::sqrt(v), sf::sqrt(x));
        }
    }
    // Differences of two binary32 values (vertex minus ray origin): any bit
    // patterns, nearby exponents (the integer path), exponent gaps around its
    // 28-step limit, equal and opposite values, zeros and subnormals.
    {
        auto fbits = [](uint32_t u) { float f; std::memcpy(&f, &u, 4); return f; };
        const float fspecial[] = {0.0f, -0.0f, 1.0f, -1.0f, INFINITY, -INFINITY, NAN, 0x1p-149f, -0x1p-149f,
                                  0x1p-126f, 0x1.fffffep127f, -0x1.fffffep127f, 3.0f, 0.1f, 1e-30f, 1e30f};
        for (float a : fspecial) for (float b : fspecial) {
            const double ref = double(a) - double(b);
            if (!same(ref, sf::sub_floats(a, b))) report("sub_floats", a, b, ref, sf::sub_floats(a, b));
        }
        for (long i = 0; i < n; ++i) {
            float a, b;
            const uint32_t r = uint32_t(g());
            switch (i % 4) {
                case 0: a = fbits(uint32_t(g())); b = fbits(uint32_t(g())); break;
                case 1: {  // exponents within 40 of each other
                    a = fbits(uint32_t(g()));
                    const int32_t ea = int32_t((sf::fbits(a) >> 23) & 0xff);
                    const int32_t eb = std::min(254, std::max(0, ea + int32_t(r % 81) - 40));
                    b = fbits((uint32_t(g()) & 0x807fffffu) | (uint32_t(eb) << 23));
                    break;
                }
                case 2: a = fbits(uint32_t(g())); b = (r & 1) ? a : -a; break;
                default: {  // neighbours and near-cancelling values
                    a = float(int32_t(r % 2000001) - 1000000) * 0.001f;
                    b = std::nextafter(a, (g() & 1) ? INFINITY : -INFINITY);
                    if (g() & 1) b = std::ldexp(b, int(g() % 61) - 30);
                    break;
                }
            }
            const double ref = double(a) - double(b);
            if (!same(ref, sf::sub_floats(a, b))) report("sub_floats", a, b, ref, sf::sub_floats(a, b));
        }
    }
    // floor / fmod(x, 2) / to_i64 (texture coordinate wrapping).
    for (long i = 0; i < n / 4; ++i) {
        double x;
        const uint64_t r = g();
        switch (i % 3) {
            case 0: x = std::ldexp(double(int64_t(r) >> 11), -int(g() % 80)); break;
            case 1: x = (double(int(g() % 4001)) - 2000) / 8.0; break;
            default: x = (double(r >> 11) * 0x1p-53 - 0.5) * std::ldexp(1.0, int(g() % 130) - 65); break;
        }
        const sf::f64 a{db(x)};
        if (!same(std::floor(x), sf::floor_(a))) report("floor", x, 0, std::floor(x), sf::floor_(a));
        if (!same(std::fmod(x, 2.0), sf::fmod2(a))) report("fmod2", x, 2, std::fmod(x, 2.0), sf::fmod2(a));
        const double f = std::floor(x);
        if (std::fabs(f) < 0x1p61 && sf::to_i64(sf::floor_(a)) != int64_t(f)) report("to_i64", x, 0, f, sf::floor_(a));
    }
    std::printf("sf64: %ld random cases, %d failures\n", n, failures);
    return failures ? 1 : 0;
}
