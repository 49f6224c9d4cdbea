// util_files.h -- file helpers of the C++ API (src/util_files.h:1-50).
#pragma once

#include <cstdio>
#include <optional>
#include <vector>

#include "field2d.h"

namespace ghcm {

// Size of an open file; the file position is undefined afterwards.
std::optional<size_t> file_size(FILE* f);

// Heat-map PNG of a time field: hue over [min, max] of the finite cells, +INF black.
bool file_write_time_field(char const* file_name, Field2D<float> const& field_time);
bool file_write_time_field(char const* file_name, Field2D<float> const& field_time, float f_min,
                           float f_max);

// Radiance HDR (RGBE) image of a field.
bool file_write_field_hdr(char const* file_name, Field2D<float> const& field_time);

template <typename T>
std::optional<std::vector<T>> file_read_all(char const* const file_name) {
    if (!file_name || !*file_name) return std::nullopt;
    FILE* f = std::fopen(file_name, "rb");
    if (!f) return std::nullopt;
    const auto size = file_size(f);
    if (!size || std::fseek(f, 0, SEEK_SET) != 0) {
        std::fclose(f);
        return std::nullopt;
    }
    std::vector<T> buf(*size / sizeof(T));
    const size_t got = std::fread(buf.data(), sizeof(T), buf.size(), f);
    std::fclose(f);
    if (got != buf.size()) return std::nullopt;
    return buf;
}

}  // namespace ghcm
