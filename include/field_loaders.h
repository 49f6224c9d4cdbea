// field_loaders.h -- image loaders of the C++ API (src/field_loaders.h:1-11).
//
// Same names and meaning as the reference: a grey image as a float field, or an RGB image
// reduced to luminosity on the GPU.  Images are PNG; the decoding and the 8-bit -> float
// mapping follow stb_image's stbi_loadf (see ghcm_io.h).  Failures return an empty
// optional (the reference prints and returns empty too).
#pragma once

#include <optional>

#include "field2d.h"

namespace ghcm {

// Direct load of a mono-channel (grey) image.
std::optional<Field2D<float>> field2d_img_monochannel(char const* file_path);

// RGB image converted to grey (perceptual luminosity) by the GPU kernel of device
// `ordinal` (the reference takes the first OpenCL GPU).
std::optional<Field2D<float>> field2d_img_rgb_luminosity(char const* file_path, int ordinal = 0);

}  // namespace ghcm
