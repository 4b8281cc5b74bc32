#pragma once
// film_host.h — host half of the film encoders (reference src/io/image.cpp:101-124,
// :201-236): input checks with the reference's messages, the display thresholds
// the device searches, and the PNG / Radiance HDR containers.
#include <cstdint>
#include <string>
#include <vector>

namespace cr {

// Display byte of the PNG writer for one channel value: exposure scale,
// filmic rational curve in binary64, sRGB in float, rounded to 8 bits.
uint8_t display_value(float value, double scale);

// t[k-1] = smallest non-negative float whose display byte is >= k (k = 1..255),
// +inf when no finite value reaches k.
void display_thresholds(double scale, float t[255]);

// Throws the reference's dimension message ("<label>: image dimensions ...").
void check_dimensions(int width, int height, const std::string& label);
// Throws the reference's exposure message.
void check_exposure(float exposure);

// Radiance header the HDR writer emits for a width x height image.
std::string hdr_header(int width, int height);

// Adler/CRC/zlib PNG container (8-bit RGB, no interlace) of rgb8; false on I/O failure.
bool write_png_rgb8(const std::string& path, int width, int height, const uint8_t* rgb8);

}  // namespace cr
