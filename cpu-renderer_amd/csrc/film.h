#pragma once
// film.h — film encode shared by the device kernels and the host: display
// bytes for the PNG writer and Radiance RGBE with its scanline run-length
// coding for the HDR writer (reference src/io/image.cpp:101-124, :201-236, and
// the RGBE and run-length rules of the pinned stb_image_write 1.16 it calls).

#include <stdint.h>

#include "sf64.h"

namespace cr {

constexpr int kDisplayLevels = 256;

// Display byte of one channel: the number of thresholds t[0..254] not above v,
// where t[k-1] is the smallest non-negative float whose display byte is >= k
// (computed on the host from the display curve, which is monotone in v).
CR_HDI uint32_t display_byte(float v, const float* t) {
    uint32_t n = 0;
#pragma unroll
    for (uint32_t step = 128; step; step >>= 1)
        if (n + step <= 255u && t[n + step - 1] <= v) n += step;
    return n;
}

// Shared-exponent RGBE: mantissas scaled by frexp(max)·256/max and truncated,
// exponent biased by 128; black below 1e-32.
CR_HDI void rgbe_of(float r, float g, float b, uint8_t out[4]) {
    const float gb = g > b ? g : b;
    const float m = r > gb ? r : gb;
    if (m < 1e-32f) {
        out[0] = out[1] = out[2] = out[3] = 0;
        return;
    }
    // m >= 1e-32 is a normal float: m = f * 2^e with f in [0.5, 1).
    const uint32_t bits = sf::fbits(m);
    const int e = int((bits >> 23) & 0xffu) - 126;
    const float f = sf::bitsf((bits & 0x807fffffu) | (126u << 23));
    const float norm = f * 256.0f / m;
    out[0] = uint8_t(r * norm);
    out[1] = uint8_t(g * norm);
    out[2] = uint8_t(b * norm);
    out[3] = uint8_t(e + 128);
}

// Largest run-length coded size of one component of a scanline of width w.
CR_HDI uint32_t rle_bound(uint32_t w) { return w + (w + 127) / 128; }

// Run-length codes w bytes c[0], c[stride], ... into out; returns the byte
// count.  Literal blocks (length byte 1..128, then the bytes) run up to the next
// three equal bytes; runs are written as (128 + length, byte) in blocks of at
// most 127.
template <class In, class Out>
CR_HDI uint32_t rle_component(In c, uint32_t stride, uint32_t w, Out out) {
    uint32_t x = 0, o = 0;
    while (x < w) {
        uint32_t r = x;
        while (r + 2 < w && !(c[r * stride] == c[(r + 1) * stride] && c[r * stride] == c[(r + 2) * stride])) ++r;
        const bool run = r + 2 < w;
        if (!run) r = w;
        while (x < r) {
            const uint32_t len = r - x < 128u ? r - x : 128u;
            out[o++] = uint8_t(len);
            for (uint32_t k = 0; k < len; ++k) out[o++] = c[(x + k) * stride];
            x += len;
        }
        if (run) {
            const uint8_t v = c[x * stride];
            while (r < w && c[r * stride] == v) ++r;
            while (x < r) {
                const uint32_t len = r - x < 127u ? r - x : 127u;
                out[o++] = uint8_t(128 + len);
                out[o++] = v;
                x += len;
            }
        }
    }
    return o;
}

}  // namespace cr
