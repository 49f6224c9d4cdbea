// png_codec.h -- PNG / zlib codec used by the field loaders and the result writers.
#pragma once

#include <cstddef>
#include <cstdint>
#include <string>
#include <vector>

namespace ghcm {

// A decoded PNG: samples widened to 16 bits, channels interleaved (1 grey, 2 grey+alpha,
// 3 RGB, 4 RGBA; palette images arrive expanded to RGB or RGBA).
struct PngImage {
    uint32_t w = 0, h = 0;
    int channels = 0;
    bool depth16 = false;            // samples span 0..65535 (else 0..255)
    std::vector<uint16_t> samples;
};

// Inflates a zlib stream, stopping once `limit` bytes are out.
bool zlib_inflate(const uint8_t* src, size_t n, std::vector<uint8_t>& out, std::string& err,
                  size_t limit = SIZE_MAX);
std::vector<uint8_t> zlib_deflate(const uint8_t* src, size_t n);
uint32_t png_crc32(const uint8_t* p, size_t n);

bool png_decode(const uint8_t* data, size_t n, PngImage& img, std::string& err);
// 8-bit pixels with `req_comp` channels (0 = as decoded), converted as stb_image does for
// a requested channel count, 16-bit samples reduced to their high byte.
std::vector<uint8_t> png_to_8bit(const PngImage& img, int req_comp);
// 8-bit PNG of `channels` (1-4) interleaved channels.
std::vector<uint8_t> png_encode(const uint8_t* pixels, uint32_t w, uint32_t h, int channels);

}  // namespace ghcm
