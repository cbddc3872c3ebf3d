// CTTypes.h — boundary types of the backend plugin API, restated with the
// same names, fields and memory layout as the reference's
// QtCTReconstruction/CTTypes.h:15-156 so that HipBackend compiles against
// either header unchanged:
//   Buffer2D   row-major data[y*width + x]
//   Buffer3D   data[(z*height + y)*width + x] plus optional axis coordinates
//   Sinogram   data is [num_angles][detector_bins] (width = bins)
//   ReconstructionParams::FilterType in the reference's order
#pragma once

#include <complex>
#include <cstddef>
#include <vector>

namespace ct {

using Complex = std::complex<float>;

struct Buffer2D {
    std::vector<float> data;
    size_t width = 0;
    size_t height = 0;

    Buffer2D() = default;
    Buffer2D(size_t w, size_t h, float value = 0.0f) : data(w * h, value), width(w), height(h) {}

    bool empty() const { return data.empty(); }
    size_t size() const { return height; }
    void assign(size_t w, size_t h, float value = 0.0f) {
        width = w;
        height = h;
        data.assign(w * h, value);
    }
    float* operator[](size_t y) { return data.data() + y * width; }
    const float* operator[](size_t y) const { return data.data() + y * width; }
    float& at(size_t x, size_t y) { return data[y * width + x]; }
    const float& at(size_t x, size_t y) const { return data[y * width + x]; }
    float* rowPtr(size_t y) { return data.data() + y * width; }
    const float* rowPtr(size_t y) const { return data.data() + y * width; }
};

using Slice = Buffer2D;

struct Buffer3D {
    std::vector<float> data;
    size_t width = 0;
    size_t height = 0;
    size_t depth = 0;
    std::vector<float> x_coords, y_coords, z_coords;

    Buffer3D() = default;
    Buffer3D(size_t w, size_t h, size_t d, float value = 0.0f)
        : data(w * h * d, value), width(w), height(h), depth(d) {}

    bool empty() const { return data.empty(); }
    void assign(size_t w, size_t h, size_t d, float value = 0.0f) {
        width = w;
        height = h;
        depth = d;
        data.assign(w * h * d, value);
    }
    float& at(size_t x, size_t y, size_t z) { return data[(z * height + y) * width + x]; }
    const float& at(size_t x, size_t y, size_t z) const { return data[(z * height + y) * width + x]; }
    Buffer2D getSlice(size_t z) const {
        if (z >= depth) return Buffer2D();
        Buffer2D s(width, height);
        const size_t n = width * height;
        std::copy(data.begin() + z * n, data.begin() + (z + 1) * n, s.data.begin());
        return s;
    }
    // A slice of another shape is dropped silently, as in the reference.
    void setSlice(size_t z, const Buffer2D& s) {
        if (z < depth && s.width == width && s.height == height)
            std::copy(s.data.begin(), s.data.end(), data.begin() + z * width * height);
    }
};

using Volume = Buffer3D;

struct Point {
    float x, y, z, hu;
};
using PointCloud = std::vector<Point>;

struct Sinogram {
    Buffer2D data;  // [num_angles][detector_bins]
    std::vector<float> angles_deg;
    float detector_spacing_mm = 1.0f;
    float original_min_hu = 0.0f;
    float original_max_hu = 0.0f;
};

struct ReconstructionParams {
    enum class FilterType { Ramp, SheppLogan, Hamming, Cosine, Hann, Bartlett };
    FilterType filter = FilterType::SheppLogan;
    size_t num_angles = 180;
    bool zero_padding = true;
};

}  // namespace ct
