// HipVolumeGenerator.cpp — IVolumeGenerator over the C ABI of libct_hip.so.
#include "HipVolumeGenerator.h"

#include "../../include/ct_hip.h"

#include <stdexcept>
#include <string>
#include <vector>

namespace ct {

namespace {
// utils::linspace (Utils.cpp:12-26): n evenly spaced values from a to b.
std::vector<float> axis(float a, float b, size_t n) {
    std::vector<float> v(n);
    if (n == 1) {
        v[0] = a;
        return v;
    }
    const float step = (b - a) / (float)(n - 1);
    for (size_t i = 0; i < n; ++i) v[i] = a + step * (float)i;
    return v;
}
}  // namespace

HipVolumeGenerator::HipVolumeGenerator(int device) {
    if (ct_hip_device_count() <= 0) return;  // isAvailable() reports false
    if (ct_hip_create(device, 0, &m_ctx) != CT_OK) m_ctx = nullptr;
}

HipVolumeGenerator::~HipVolumeGenerator() { ct_hip_destroy(m_ctx); }

bool HipVolumeGenerator::isAvailable() const { return m_ctx != nullptr; }

Volume HipVolumeGenerator::generateBrainHU(const Generator3D::Params& params) {
    Volume vol;
    const size_t w = params.shape[0], h = params.shape[1], d = params.shape[2];
    if (w == 0 || h == 0 || d == 0) return vol;
    if (!m_ctx) throw std::runtime_error("HipVolumeGenerator: no HIP device");
    vol.assign(w, h, d, 0.0f);
    const int n = (int)params.num_ellipsoids;
    std::vector<ct_ellipsoid> e((size_t)n > 0 ? (size_t)n : 1);
    if (ct_phantom_params(params.seed, n, e.data()) != CT_OK)
        throw std::runtime_error("HipVolumeGenerator: bad phantom parameters");
    const int rc = ct_hip_phantom(m_ctx, e.data(), n, (int)w, (int)h, (int)d, vol.data.data());
    if (rc != CT_OK)
        throw std::runtime_error("HipVolumeGenerator::generateBrainHU failed (" +
                                 std::to_string(rc) + "): " + ct_hip_last_error(m_ctx));
    const float R = params.brain_radius_mm;
    vol.x_coords = axis(-R, R, w);
    vol.y_coords = axis(-R, R, h);
    vol.z_coords = axis(-R, R, d);
    return vol;
}

}  // namespace ct
