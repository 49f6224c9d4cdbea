// field2d.h -- row-major 2-D field and index helpers of the C++ API.
//
// Same shape and semantics as the reference's Field2D (src/util_field.h:1-58) and
// dim_* helpers (src/util_dim.h:1-29): W x H cells, idx = x + y*W, get_or() returns
// the default outside the grid.  glm's vector types are replaced by two small
// structs with the members the API uses (x, y).
#pragma once

#include <cassert>
#include <cstddef>
#include <cstdint>
#include <vector>

namespace ghcm {

struct uint2 {
    uint32_t x = 0, y = 0;
    constexpr uint2() = default;
    constexpr uint2(uint32_t a, uint32_t b) : x(a), y(b) {}
    constexpr explicit uint2(uint32_t a) : x(a), y(a) {}
    constexpr bool operator==(const uint2& o) const { return x == o.x && y == o.y; }
    constexpr bool operator!=(const uint2& o) const { return !(*this == o); }
};

struct int2 {
    int32_t x = 0, y = 0;
    constexpr int2() = default;
    constexpr int2(int32_t a, int32_t b) : x(a), y(b) {}
    constexpr explicit int2(uint2 v) : x((int32_t)v.x), y((int32_t)v.y) {}
    constexpr int2 operator+(const int2& o) const { return {x + o.x, y + o.y}; }
    constexpr int2 operator-(const int2& o) const { return {x - o.x, y - o.y}; }
};

inline bool dim_in_bounds(uint2 sz, uint2 i) { return i.x < sz.x && i.y < sz.y; }
inline bool dim_in_bounds(uint2 sz, int2 i) {
    return i.x >= 0 && i.y >= 0 && (uint32_t)i.x < sz.x && (uint32_t)i.y < sz.y;
}
inline size_t dim_size(uint2 sz) { return (size_t)sz.x * sz.y; }
inline size_t dim_idx(uint2 sz, uint2 i) { return (size_t)i.x + (size_t)i.y * sz.x; }
inline uint2 dim_pos(uint2 sz, size_t i) { return {(uint32_t)(i % sz.x), (uint32_t)(i / sz.x)}; }

template <typename T>
class Field2D {
    std::vector<T> data_;
    uint2 size_{0, 0};

public:
    Field2D() = default;
    explicit Field2D(uint2 sz) : data_(dim_size(sz)), size_(sz) {}
    Field2D(uint2 sz, T v) : data_(dim_size(sz), v), size_(sz) {}
    Field2D(uint2 sz, const T* values) : data_(values, values + dim_size(sz)), size_(sz) {}

    size_t size_flat() const { return dim_size(size_); }
    uint2 size() const { return size_; }

    bool in_bounds(uint2 p) const { return dim_in_bounds(size_, p); }
    bool in_bounds(int2 p) const { return dim_in_bounds(size_, p); }

    T get_or(const T& dflt, int2 p) const { return in_bounds(p) ? data_[dim_idx(size_, uint2((uint32_t)p.x, (uint32_t)p.y))] : dflt; }

    std::vector<T>& backing_store() { return data_; }
    const std::vector<T>& backing_store() const { return data_; }

    T& operator[](size_t i) { assert(i < data_.size()); return data_[i]; }
    const T& operator[](size_t i) const { assert(i < data_.size()); return data_[i]; }
    T& operator[](uint2 p) { return (*this)[dim_idx(size_, p)]; }
    const T& operator[](uint2 p) const { return (*this)[dim_idx(size_, p)]; }
    T& operator[](int2 p) { assert(in_bounds(p)); return (*this)[uint2((uint32_t)p.x, (uint32_t)p.y)]; }
    const T& operator[](int2 p) const { assert(in_bounds(p)); return (*this)[uint2((uint32_t)p.x, (uint32_t)p.y)]; }
};

}  // namespace ghcm
