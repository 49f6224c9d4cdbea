/*
 * ghcm_io.h -- C-ABI of the field loaders, result writers and test runner (libghcm.so).
 *
 * Replaces the reference's image / file surface:
 *   field2d_img_monochannel, field2d_img_rgb_luminosity   src/field_loaders.h:6-11
 *   img_proc_rgb_luminosity (GPU kernel)                  data/img_proc_rgb_luminosity.cl:5-19
 *   file_write_time_field, file_write_field_hdr           src/util_files.h:9-19
 *   the .blob / .csv / .png outputs of test_run            src/test_runner.cpp:163-217
 * Images are PNG (decoded by the library's own codec; the reference uses stb_image).
 * Every function returns a ghcm_status; messages are in ghcm_last_error().
 */
#ifndef GHCM_IO_H
#define GHCM_IO_H

#include <stddef.h>
#include <stdint.h>

#include "ghcm.h"

#ifdef __cplusplus
extern "C" {
#endif

/* A W x H image or field of `comp` interleaved float channels, allocated by the library. */
typedef struct ghcm_image {
    uint32_t w, h;
    uint32_t comp;
    uint32_t reserved;
    float*   data;
} ghcm_image;

void ghcm_image_free(ghcm_image* img);

/* Decodes a PNG file to floats the way stbi_loadf(path, req_comp) does (req_comp 1-4;
 * src/field_loaders.cpp:26-48): channels converted at the source precision (grey =
 * (77 r + 150 g + 29 b) >> 8), 16-bit samples reduced to their high byte, then every
 * colour channel mapped v -> (v / 255)^2.2 and alpha v -> v / 255. */
int ghcm_image_load_f32(const char* path, int req_comp, ghcm_image* out);

/* Decodes PNG bytes to 8-bit pixels with `req_comp` channels (0 = as stored).  Call with
 * pixels = NULL to get *w, *h and *comp; then with a buffer of w*h*comp bytes. */
int ghcm_png_decode8(const void* bytes, size_t n, int req_comp, uint32_t* w, uint32_t* h,
                     uint32_t* comp, uint8_t* pixels, size_t capacity);

/* Writes an 8-bit PNG of 1-4 interleaved channels. */
int ghcm_png_write8(const char* path, const uint8_t* pixels, uint32_t w, uint32_t h, uint32_t comp);

/* field2d_img_monochannel (src/field_loaders.cpp:53-59): grey image as a float field. */
int ghcm_field_load_mono(const char* path, ghcm_image* out);

/* field2d_img_rgb_luminosity (src/field_loaders.cpp:62-102): RGB image, luminosity
 * formed on the GPU of `ctx`. */
int ghcm_field_load_rgb_luminosity(ghcm_ctx* ctx, const char* path, ghcm_image* out);

/* The luminosity kernel on device buffers: d_rgb holds w*h interleaved RGB floats,
 * d_out receives w*h floats (data/img_proc_rgb_luminosity.cl:5-19). */
int ghcm_rgb_luminosity_device(ghcm_ctx* ctx, const float* d_rgb, uint32_t w, uint32_t h,
                               float* d_out);

/* file_write_time_field (src/util_files.cpp:52-97): heat-map PNG of a time field, hue
 * (5/6)(t - f_min)/(f_max - f_min) per finite cell, black for +INF.  auto_range != 0
 * takes f_min / f_max from the finite cells. */
int ghcm_write_time_field_png(const char* path, const float* field, uint32_t w, uint32_t h,
                              float f_min, float f_max, int auto_range);

/* file_write_field_hdr (src/util_files.cpp:40-50): Radiance RGBE image of a grey field. */
int ghcm_write_field_hdr(const char* path, const float* field, uint32_t w, uint32_t h);

/* The result blob of test_run (src/test_runner.cpp:194-200): two size_t (W, H) then W*H
 * floats, host byte order. */
int ghcm_blob_write(const char* path, const float* field, uint32_t w, uint32_t h);
int ghcm_blob_read(const char* path, ghcm_image* out);

/* test_run (src/test_runner.cpp:163-217) for one job: `trials` solves of `cost` with the
 * named solver ("fmm", "fim" or "hcm"; "hcm" uses ctx's configuration), then
 * <out_dir>/<solver_name>_<job_name>.{blob,png,csv}.  The solve is skipped when the CSV
 * exists and is not empty, as in the reference. */
int ghcm_test_run(const char* solver, ghcm_ctx* ctx, const char* solver_name, const char* job_name,
                  const float* cost, uint32_t w, uint32_t h, const uint32_t* seeds_xy,
                  uint32_t n_seeds, const char* out_dir, uint32_t trials,
                  float fim_minimum_change);

/* Parametric fields of the harness (src/test_runner.cpp:42-76), W*H floats out. */
int ghcm_field_cost_sin_checkerboard(uint32_t w, uint32_t h, float speed_amplitude,
                                     float checkerboard_res, float* out);
int ghcm_field_cost_random(uint32_t w, uint32_t h, uint64_t seed, float* out);

#ifdef __cplusplus
}
#endif
#endif
