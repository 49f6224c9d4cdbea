// field_io.cpp -- field loaders and result writers (C-ABI of ghcm_io.h, C++ of
// field_loaders.h / util_files.h).
//
// Image input follows stb_image's `stbi_loadf` as the reference calls it
// (src/field_loaders.cpp:26-48): the PNG is decoded (png_codec.cpp), converted to the
// requested channel count, and every 8-bit colour value v becomes
// (float)pow((double)(v / 255.0f), (double)2.2f), alpha v / 255.0f -- stb's LDR-to-float
// mapping with its default gamma 2.2 and scale 1.  That mapping cannot be checked against
// stb here (the library is absent), so its parity is unpinned; the pixel decoding is
// checked against an independent PNG decoder in the tests.
#include <cmath>
#include <cstdio>
#include <cstring>
#include <memory>
#include <new>
#include <string>
#include <vector>

#include "capi_common.h"
#include "field_loaders.h"
#include "ghcm.h"
#include "ghcm_io.h"
#include "png_codec.h"
#include "util_files.h"

using ghcm_capi::fail;

namespace {

bool read_file(const char* path, std::vector<uint8_t>& out) {
    if (!path) return false;
    FILE* f = std::fopen(path, "rb");
    if (!f) return false;
    bool ok = std::fseek(f, 0, SEEK_END) == 0;
    const long len = ok ? std::ftell(f) : -1;
    ok = ok && len >= 0 && std::fseek(f, 0, SEEK_SET) == 0;
    if (ok) {
        out.resize((size_t)len);
        ok = std::fread(out.data(), 1, out.size(), f) == out.size();
    }
    std::fclose(f);
    return ok;
}

bool write_file(const char* path, const void* data, size_t n) {
    FILE* f = std::fopen(path, "wb");
    if (!f) return false;
    const bool ok = std::fwrite(data, 1, n, f) == n;
    return std::fclose(f) == 0 && ok;
}

// stb's LDR -> float mapping for 8-bit values, one table entry per value.
struct GammaTable {
    float lut[256];
    GammaTable() {
        for (int v = 0; v < 256; ++v)
            lut[v] = (float)(std::pow((double)((float)v / 255.0f), (double)2.2f) * (double)1.0f);
    }
};
const GammaTable& gamma_table() {
    static const GammaTable t;
    return t;
}

int load_f32(const char* path, int req_comp, ghcm_image* out) {
    return ghcm_capi::no_throw([&]() -> int {
        if (!path || !out) return fail(GHCM_ERR_INVALID_ARG, "null argument");
        std::memset(out, 0, sizeof *out);
        if (req_comp < 1 || req_comp > 4) return fail(GHCM_ERR_INVALID_ARG, "req_comp must be 1..4");
        std::vector<uint8_t> bytes;
        if (!read_file(path, bytes)) return fail(GHCM_ERR_INVALID_ARG, "cannot read image file %s", path);
        ghcm::PngImage img;
        std::string err;
        if (!ghcm::png_decode(bytes.data(), bytes.size(), img, err))
            return fail(GHCM_ERR_INVALID_ARG, "cannot decode %s: %s", path, err.c_str());
        const std::vector<uint8_t> px = ghcm::png_to_8bit(img, req_comp);
        float* data = static_cast<float*>(std::malloc(px.size() * sizeof(float)));
        if (!data) return fail(GHCM_ERR_OUT_OF_MEMORY, "image of %zu floats", px.size());
        const float* lut = gamma_table().lut;
        // Alpha (the last channel of 2 or 4) is linear.
        const bool has_alpha = req_comp == 2 || req_comp == 4;
        for (size_t i = 0; i < px.size(); ++i) {
            const bool alpha = has_alpha && (i % req_comp) == (size_t)(req_comp - 1);
            data[i] = alpha ? (float)px[i] / 255.0f : lut[px[i]];
        }
        out->w = img.w;
        out->h = img.h;
        out->comp = (uint32_t)req_comp;
        out->data = data;
        return GHCM_OK;
    });
}

// Hue of the reference's heat map (src/util_files.cpp:13-27): h in [0, 1] -> RGB8.
void hue_to_rgb(float h, uint8_t* o) {
    const float s = std::fmod(h, 1.0f) * 6.0f;
    const uint8_t x = (uint8_t)((1.0f - std::fabs(std::fmod(s, 2.0f) - 1.0f)) * 255.0f);
    switch ((int)s) {
        case 0: o[0] = 255; o[1] = x;   o[2] = 0;   break;
        case 1: o[0] = x;   o[1] = 255; o[2] = 0;   break;
        case 2: o[0] = 0;   o[1] = 255; o[2] = x;   break;
        case 3: o[0] = 0;   o[1] = x;   o[2] = 255; break;
        case 4: o[0] = x;   o[1] = 0;   o[2] = 255; break;
        default: o[0] = 255; o[1] = 0;  o[2] = x;   break;
    }
}

// Radiance RGBE of one grey value (Ward's shared-exponent format).
void to_rgbe(float v, uint8_t* o) {
    if (!(v >= 1e-32f) || std::isinf(v)) {
        o[0] = o[1] = o[2] = o[3] = 0;
        return;
    }
    int e = 0;
    const float m = std::frexp(v, &e) * 256.0f / v;
    const uint8_t c = (uint8_t)(v * m);
    o[0] = o[1] = o[2] = c;
    o[3] = (uint8_t)(e + 128);
}

// One scanline component in Radiance run-length form: runs of 3 or more equal bytes as
// (128 + n, value), everything else as literal dumps (n, bytes...), n <= 127 / 128.
void rle_component(const uint8_t* v, int w, std::vector<uint8_t>& out) {
    int i = 0;
    while (i < w) {
        int run = 1;
        while (i + run < w && run < 127 && v[i + run] == v[i]) ++run;
        if (run >= 3) {
            out.push_back((uint8_t)(128 + run));
            out.push_back(v[i]);
            i += run;
            continue;
        }
        // literal dump up to the next run of 3
        int j = i, n = 0;
        while (j < w && n < 128) {
            if (j + 2 < w && v[j] == v[j + 1] && v[j] == v[j + 2]) break;
            ++j;
            ++n;
        }
        out.push_back((uint8_t)n);
        out.insert(out.end(), v + i, v + i + n);
        i += n;
    }
}

}  // namespace

// ---------------------------------------------------------------------------------------
// C-ABI
// ---------------------------------------------------------------------------------------

extern "C" void ghcm_image_free(ghcm_image* img) {
    if (!img) return;
    std::free(img->data);
    std::memset(img, 0, sizeof *img);
}

extern "C" int ghcm_image_load_f32(const char* path, int req_comp, ghcm_image* out) {
    return load_f32(path, req_comp, out);
}

extern "C" int ghcm_png_decode8(const void* bytes, size_t n, int req_comp, uint32_t* w, uint32_t* h,
                                uint32_t* comp, uint8_t* pixels, size_t capacity) {
    return ghcm_capi::no_throw([&]() -> int {
        if (!bytes || !w || !h || !comp) return fail(GHCM_ERR_INVALID_ARG, "null argument");
        if (req_comp < 0 || req_comp > 4) return fail(GHCM_ERR_INVALID_ARG, "req_comp must be 0..4");
        ghcm::PngImage img;
        std::string err;
        if (!ghcm::png_decode(static_cast<const uint8_t*>(bytes), n, img, err))
            return fail(GHCM_ERR_INVALID_ARG, "PNG: %s", err.c_str());
        *w = img.w;
        *h = img.h;
        *comp = (uint32_t)(req_comp ? req_comp : img.channels);
        if (!pixels) return GHCM_OK;
        const std::vector<uint8_t> px = ghcm::png_to_8bit(img, req_comp);
        if (capacity < px.size()) return fail(GHCM_ERR_INVALID_ARG, "buffer of %zu bytes, need %zu", capacity, px.size());
        std::memcpy(pixels, px.data(), px.size());
        return GHCM_OK;
    });
}

extern "C" int ghcm_png_write8(const char* path, const uint8_t* pixels, uint32_t w, uint32_t h,
                               uint32_t comp) {
    return ghcm_capi::no_throw([&]() -> int {
        if (!path || !pixels || !w || !h || comp < 1 || comp > 4)
            return fail(GHCM_ERR_INVALID_ARG, "bad PNG write arguments");
        const std::vector<uint8_t> png = ghcm::png_encode(pixels, w, h, (int)comp);
        if (!write_file(path, png.data(), png.size())) return fail(GHCM_ERR_INVALID_ARG, "cannot write %s", path);
        return GHCM_OK;
    });
}

extern "C" int ghcm_field_load_mono(const char* path, ghcm_image* out) {
    return load_f32(path, 1, out);
}

extern "C" int ghcm_field_load_rgb_luminosity(ghcm_ctx* ctx, const char* path, ghcm_image* out) {
    return ghcm_capi::no_throw([&]() -> int {
        if (!ctx) return fail(GHCM_ERR_INVALID_ARG, "ctx is null");
        ghcm_image rgb;
        int rc = load_f32(path, 3, &rgb);
        if (rc) return rc;
        std::unique_ptr<ghcm_image, void (*)(ghcm_image*)> guard(&rgb, ghcm_image_free);
        const size_t n = (size_t)rgb.w * rgb.h;
        void *d_in = nullptr, *d_out = nullptr;
        if ((rc = ghcm_dev_alloc(ctx, n * 3 * sizeof(float), &d_in))) return rc;
        if ((rc = ghcm_dev_alloc(ctx, n * sizeof(float), &d_out))) {
            ghcm_dev_free(ctx, d_in);
            return rc;
        }
        float* lum = static_cast<float*>(std::malloc(n * sizeof(float)));
        if (!lum) rc = fail(GHCM_ERR_OUT_OF_MEMORY, "image of %zu floats", n);
        if (!rc) rc = ghcm_dev_upload(ctx, d_in, rgb.data, n * 3 * sizeof(float));
        if (!rc) rc = ghcm_rgb_luminosity_device(ctx, static_cast<float*>(d_in), rgb.w, rgb.h, static_cast<float*>(d_out));
        if (!rc) rc = ghcm_dev_download(ctx, lum, d_out, n * sizeof(float));
        ghcm_dev_free(ctx, d_in);
        ghcm_dev_free(ctx, d_out);
        if (rc) {
            std::free(lum);
            return rc;
        }
        std::memset(out, 0, sizeof *out);
        out->w = rgb.w;
        out->h = rgb.h;
        out->comp = 1;
        out->data = lum;
        return GHCM_OK;
    });
}

extern "C" int ghcm_write_time_field_png(const char* path, const float* field, uint32_t w, uint32_t h,
                                         float f_min, float f_max, int auto_range) {
    return ghcm_capi::no_throw([&]() -> int {
        if (!path || !field || !w || !h) return fail(GHCM_ERR_INVALID_ARG, "bad time-field write arguments");
        const size_t n = (size_t)w * h;
        if (auto_range) {
            f_min = INFINITY;
            f_max = -INFINITY;
            for (size_t i = 0; i < n; ++i)
                if (!std::isinf(field[i])) {
                    f_min = std::fmin(f_min, field[i]);
                    f_max = std::fmax(f_max, field[i]);
                }
        }
        std::vector<uint8_t> px(n * 3, 0);
        if (!(f_max < f_min)) {
            const float range = (f_min == f_max) ? 1.0f : (f_max - f_min);
            for (size_t i = 0; i < n; ++i)
                if (!std::isinf(field[i])) hue_to_rgb((5.0f / 6.0f) * (field[i] - f_min) / range, &px[3 * i]);
        }
        return ghcm_png_write8(path, px.data(), w, h, 3);
    });
}

extern "C" int ghcm_write_field_hdr(const char* path, const float* field, uint32_t w, uint32_t h) {
    return ghcm_capi::no_throw([&]() -> int {
        if (!path || !field || !w || !h) return fail(GHCM_ERR_INVALID_ARG, "bad HDR write arguments");
        std::vector<uint8_t> out;
        char head[128];
        const int hl = std::snprintf(head, sizeof head,
                                     "#?RADIANCE\n# ghcm time field\nFORMAT=32-bit_rle_rgbe\n\n-Y %u +X %u\n", h, w);
        out.insert(out.end(), head, head + hl);
        std::vector<uint8_t> px((size_t)w * 4), comp(w);
        for (uint32_t y = 0; y < h; ++y) {
            for (uint32_t x = 0; x < w; ++x) to_rgbe(field[(size_t)y * w + x], &px[4 * (size_t)x]);
            if (w < 8 || w >= 32768) {                       // flat scanline
                out.insert(out.end(), px.begin(), px.end());
                continue;
            }
            out.push_back(2);
            out.push_back(2);
            out.push_back((uint8_t)(w >> 8));
            out.push_back((uint8_t)(w & 255));
            for (int c = 0; c < 4; ++c) {
                for (uint32_t x = 0; x < w; ++x) comp[x] = px[4 * (size_t)x + c];
                rle_component(comp.data(), (int)w, out);
            }
        }
        if (!write_file(path, out.data(), out.size())) return fail(GHCM_ERR_INVALID_ARG, "cannot write %s", path);
        return GHCM_OK;
    });
}

extern "C" int ghcm_blob_write(const char* path, const float* field, uint32_t w, uint32_t h) {
    return ghcm_capi::no_throw([&]() -> int {
        if (!path || !field) return fail(GHCM_ERR_INVALID_ARG, "null argument");
        FILE* f = std::fopen(path, "wb");
        if (!f) return fail(GHCM_ERR_INVALID_ARG, "cannot write %s", path);
        const size_t x = w, y = h;
        bool ok = std::fwrite(&x, sizeof x, 1, f) == 1 && std::fwrite(&y, sizeof y, 1, f) == 1 &&
                  std::fwrite(field, sizeof(float), x * y, f) == x * y;
        ok = (std::fclose(f) == 0) && ok;
        return ok ? GHCM_OK : fail(GHCM_ERR_INVALID_ARG, "short write to %s", path);
    });
}

extern "C" int ghcm_blob_read(const char* path, ghcm_image* out) {
    return ghcm_capi::no_throw([&]() -> int {
        if (!path || !out) return fail(GHCM_ERR_INVALID_ARG, "null argument");
        std::memset(out, 0, sizeof *out);
        std::vector<uint8_t> bytes;
        if (!read_file(path, bytes)) return fail(GHCM_ERR_INVALID_ARG, "cannot read %s", path);
        size_t x = 0, y = 0;
        if (bytes.size() < 2 * sizeof(size_t)) return fail(GHCM_ERR_INVALID_ARG, "%s is too short", path);
        std::memcpy(&x, bytes.data(), sizeof x);
        std::memcpy(&y, bytes.data() + sizeof x, sizeof y);
        // x * y * 4 must not wrap: a wrapped product could match a short file.
        if (x > 0xFFFFFFFFull || y > 0xFFFFFFFFull || (y && x > (SIZE_MAX / sizeof(float)) / y) ||
            bytes.size() != 2 * sizeof(size_t) + x * y * sizeof(float))
            return fail(GHCM_ERR_INVALID_ARG, "%s: size header does not match the file", path);
        float* data = static_cast<float*>(std::malloc(std::max<size_t>(1, x * y) * sizeof(float)));
        if (!data) return fail(GHCM_ERR_OUT_OF_MEMORY, "blob of %zu floats", x * y);
        std::memcpy(data, bytes.data() + 2 * sizeof(size_t), x * y * sizeof(float));
        out->w = (uint32_t)x;
        out->h = (uint32_t)y;
        out->comp = 1;
        out->data = data;
        return GHCM_OK;
    });
}

// ---------------------------------------------------------------------------------------
// C++ API
// ---------------------------------------------------------------------------------------
namespace ghcm {

namespace {
std::optional<Field2D<float>> to_field(int rc, ghcm_image& img) {
    if (rc != GHCM_OK) return std::nullopt;
    Field2D<float> f(uint2(img.w, img.h), img.data);
    ghcm_image_free(&img);
    return f;
}
}  // namespace

std::optional<Field2D<float>> field2d_img_monochannel(char const* file_path) {
    ghcm_image img{};
    const int rc = ghcm_field_load_mono(file_path, &img);
    return to_field(rc, img);
}

std::optional<Field2D<float>> field2d_img_rgb_luminosity(char const* file_path, int ordinal) {
    ghcm_ctx* ctx = nullptr;
    if (ghcm_ctx_create(ordinal, nullptr, &ctx) != GHCM_OK) return std::nullopt;
    ghcm_image img{};
    const int rc = ghcm_field_load_rgb_luminosity(ctx, file_path, &img);
    ghcm_ctx_destroy(ctx);
    return to_field(rc, img);
}

std::optional<size_t> file_size(FILE* const f) {
    if (!f || std::fseek(f, 0, SEEK_END) != 0) return std::nullopt;
    const long len = std::ftell(f);
    if (len < 0) return std::nullopt;
    return (size_t)len;
}

bool file_write_time_field(char const* file_name, Field2D<float> const& field_time) {
    return ghcm_write_time_field_png(file_name, field_time.backing_store().data(), field_time.size().x,
                                     field_time.size().y, 0.f, 0.f, 1) == GHCM_OK;
}

bool file_write_time_field(char const* file_name, Field2D<float> const& field_time, float f_min,
                           float f_max) {
    return ghcm_write_time_field_png(file_name, field_time.backing_store().data(), field_time.size().x,
                                     field_time.size().y, f_min, f_max, 0) == GHCM_OK;
}

bool file_write_field_hdr(char const* file_name, Field2D<float> const& field_time) {
    return ghcm_write_field_hdr(file_name, field_time.backing_store().data(), field_time.size().x,
                                field_time.size().y) == GHCM_OK;
}

}  // namespace ghcm
