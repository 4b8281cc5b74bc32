// film_host.cpp — host half of the film encoders.
#include "film_host.h"

#include <zlib.h>

#include <algorithm>
#include <cmath>
#include <cstdio>
#include <cstring>
#include <limits>
#include <stdexcept>

namespace cr {

namespace {
float srgb_encode(float v) { return v <= 0.0031308f ? 12.92f * v : 1.055f * std::pow(v, 1 / 2.4f) - 0.055f; }
float float_of(uint32_t u) {
    float f;
    std::memcpy(&f, &u, 4);
    return f;
}
}  // namespace

// reference src/io/image.cpp:117-124 (displayChannel) and src/math.h:62 (linearToSrgb)
uint8_t display_value(float value, double scale) {
    const double x = double(value) * scale;
    const double filmic = std::clamp((x * (2.51 * x + 0.03)) / (x * (2.43 * x + 0.59) + 0.14), 0.0, 1.0);
    const float s = srgb_encode(float(filmic));
    return uint8_t(std::lround(std::clamp(s, 0.0f, 1.0f) * 255.0f));
}

// Bisection over the bit patterns of the non-negative floats (ordered like
// their values): the first pattern whose display byte reaches k.
void display_thresholds(double scale, float t[255]) {
    const uint32_t top = 0x7f7fffffu;  // FLT_MAX
    for (int k = 1; k <= 255; ++k) {
        if (display_value(float_of(top), scale) < k) {
            t[k - 1] = std::numeric_limits<float>::infinity();
            continue;
        }
        uint32_t lo = 0, hi = top;  // display(hi) >= k
        while (lo < hi) {
            const uint32_t mid = lo + (hi - lo) / 2;
            if (display_value(float_of(mid), scale) >= k) hi = mid;
            else lo = mid + 1;
        }
        t[k - 1] = float_of(lo);
    }
}

// reference src/io/file.h:84-89 (imagePixels)
void check_dimensions(int width, int height, const std::string& label) {
    if (width <= 0 || height <= 0 || width > 16384 || height > 16384 ||
        size_t(width) > (size_t(16) * 1024 * 1024) / size_t(height))
        throw std::runtime_error(label + ": image dimensions must be positive, at most 16384 per axis and 16777216 pixels");
}

// reference src/io/image.cpp:205-206
void check_exposure(float exposure) {
    if (!std::isfinite(exposure) || exposure < -100 || exposure > 100)
        throw std::runtime_error("PNG exposure must be finite and between -100 and 100 stops");
}

std::string hdr_header(int width, int height) {
    return "#?RADIANCE\n# Written by stb_image_write.h\nFORMAT=32-bit_rle_rgbe\n"
           "EXPOSURE=          1.0000000000000\n\n-Y " + std::to_string(height) + " +X " + std::to_string(width) + "\n";
}

namespace {
void put_be32(std::vector<uint8_t>& v, uint32_t x) {
    for (int s = 24; s >= 0; s -= 8) v.push_back(uint8_t(x >> s));
}
void chunk(std::vector<uint8_t>& out, const char* type, const uint8_t* data, size_t n) {
    put_be32(out, uint32_t(n));
    const size_t start = out.size();
    out.insert(out.end(), type, type + 4);
    out.insert(out.end(), data, data + n);
    put_be32(out, uint32_t(crc32(0, out.data() + start, uInt(n + 4))));
}
}  // namespace

bool write_png_rgb8(const std::string& path, int width, int height, const uint8_t* rgb8) {
    // Scanlines with filter type 0, deflated by zlib.
    const size_t row = size_t(width) * 3;
    std::vector<uint8_t> raw((row + 1) * size_t(height));
    for (int y = 0; y < height; ++y) {
        raw[(row + 1) * size_t(y)] = 0;
        std::memcpy(&raw[(row + 1) * size_t(y) + 1], rgb8 + row * size_t(y), row);
    }
    uLongf zn = compressBound(uLong(raw.size()));
    std::vector<uint8_t> z(zn);
    if (compress2(z.data(), &zn, raw.data(), uLong(raw.size()), 1) != Z_OK) return false;  // fastest level
    std::vector<uint8_t> out = {0x89, 'P', 'N', 'G', '\r', '\n', 0x1a, '\n'};
    uint8_t ihdr[13];
    for (int i = 0; i < 4; ++i) {
        ihdr[i] = uint8_t(uint32_t(width) >> (24 - 8 * i));
        ihdr[4 + i] = uint8_t(uint32_t(height) >> (24 - 8 * i));
    }
    ihdr[8] = 8;   // bit depth
    ihdr[9] = 2;   // RGB
    ihdr[10] = 0;  // deflate
    ihdr[11] = 0;  // adaptive filtering, filter 0 on every row
    ihdr[12] = 0;  // no interlace
    chunk(out, "IHDR", ihdr, 13);
    chunk(out, "IDAT", z.data(), zn);
    chunk(out, "IEND", nullptr, 0);
    std::FILE* f = std::fopen(path.c_str(), "wb");
    if (!f) return false;
    const bool ok = std::fwrite(out.data(), 1, out.size(), f) == out.size();
    return (std::fclose(f) == 0) && ok;
}

}  // namespace cr
