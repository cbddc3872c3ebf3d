// HipBackend.cpp — IReconstructionBackend over the C ABI of libct_hip.so.
// Semantics follow SerialBackend (SerialBackend.cpp:59-346); errors throw
// std::runtime_error, which the reference's controller catches around every
// backend call (CtReconstructionController.cpp:1145-1153, 1001-1007).
#include "HipBackend.h"

#include "../../include/ct_hip.h"

#include <stdexcept>
#include <string>

namespace ct {

HipBackend::HipBackend(int device, size_t hbm_budget_bytes) {
    if (ct_hip_device_count() <= 0) return;  // isAvailable() reports false
    if (ct_hip_create(device, hbm_budget_bytes, &m_ctx) != CT_OK) m_ctx = nullptr;
}

HipBackend::~HipBackend() { ct_hip_destroy(m_ctx); }

bool HipBackend::isAvailable() const { return m_ctx != nullptr; }

void HipBackend::check(int rc, const char* what) const {
    if (rc == CT_OK) return;
    std::string msg = std::string("HipBackend::") + what + " failed (" + std::to_string(rc) + ")";
    if (m_ctx) msg += ": " + std::string(ct_hip_last_error(m_ctx));
    throw std::runtime_error(msg);
}

Sinogram HipBackend::computeSinogram(const Buffer2D& slice, size_t num_angles,
                                     size_t detector_bins, bool /*use_parallel*/) {
    Sinogram s;
    if (slice.empty() || num_angles == 0 || detector_bins == 0) return s;  // SerialBackend.cpp:64
    if (!m_ctx) check(CT_ERR_NODEVICE, "computeSinogram");
    s.data.assign(detector_bins, num_angles, 0.0f);
    float mm[2] = {0.0f, 0.0f};
    check(ct_hip_compute_sinogram(m_ctx, slice.data.data(), (int)slice.width, (int)slice.height,
                                  (int)num_angles, (int)detector_bins, s.data.data.data(), mm),
          "computeSinogram");
    s.angles_deg.resize(num_angles);
    const float step = 180.0f / (float)num_angles;
    for (size_t a = 0; a < num_angles; ++a) s.angles_deg[a] = step * (float)a;
    s.detector_spacing_mm = 1.0f;
    s.original_min_hu = mm[0];
    s.original_max_hu = mm[1];
    return s;
}

Buffer2D HipBackend::reconstructSlice(const Sinogram& sino, size_t output_size,
                                      const ReconstructionParams& params) {
    const size_t bins = sino.data.width, na = sino.data.height;
    if (bins == 0 || na == 0 || output_size == 0) return {};  // SerialBackend.cpp:148
    if (!m_ctx) check(CT_ERR_NODEVICE, "reconstructSlice");
    if (sino.angles_deg.size() < na) throw std::invalid_argument("reconstructSlice: angles_deg too short");
    Buffer2D out(output_size, output_size, 0.0f);
    check(ct_hip_reconstruct_slice(m_ctx, sino.data.data.data(), (int)na, (int)bins,
                                   sino.angles_deg.data(), (int)output_size, (int)params.filter,
                                   out.data.data()),
          "reconstructSlice");
    return out;
}

void HipBackend::reconstructVolume(const Volume& in, Volume& out, const ReconstructionParams& params,
                                   std::function<void(int, const Buffer2D&)> onSliceDone) {
    if (in.empty()) return;
    if (!m_ctx) check(CT_ERR_NODEVICE, "reconstructVolume");
    out.assign(in.width, in.height, in.depth, 0.0f);
    out.x_coords = in.x_coords;
    out.y_coords = in.y_coords;
    out.z_coords = in.z_coords;
    check(ct_hip_reconstruct_volume(m_ctx, in.data.data(), (int)in.width, (int)in.height,
                                    (int)in.depth, (int)params.num_angles, (int)params.filter,
                                    out.data.data(), nullptr),
          "reconstructVolume");
    double t[4];
    ct_hip_last_timings(m_ctx, t);
    m_lastSinogramTimeMs = t[0];
    m_lastReconstructionTimeMs = t[1] + t[2];
    if (onSliceDone)
        for (size_t z = 0; z < out.depth; ++z) onSliceDone((int)z, out.getSlice(z));
}

void HipBackend::reconstructVolumeFromSinograms(const Volume& sinos,
                                                const std::vector<float>& angles_deg, Volume& out,
                                                const ReconstructionParams& params,
                                                std::function<void(int, const Buffer2D&)> onSliceDone) {
    const size_t nz = sinos.depth, na = sinos.height, bins = sinos.width;
    if (nz == 0 || na == 0 || bins == 0) return;
    if (!m_ctx) check(CT_ERR_NODEVICE, "reconstructVolumeFromSinograms");
    if (angles_deg.size() < na) throw std::invalid_argument("angles_deg too short");
    out.assign(bins, bins, nz, 0.0f);
    check(ct_hip_reconstruct_from_sinograms(m_ctx, sinos.data.data(), (int)nz, (int)na, (int)bins,
                                            angles_deg.data(), (int)params.filter,
                                            out.data.data()),
          "reconstructVolumeFromSinograms");
    double t[4];
    ct_hip_last_timings(m_ctx, t);
    m_lastSinogramTimeMs = 0.0;
    m_lastReconstructionTimeMs = t[1] + t[2];
    if (onSliceDone)
        for (size_t z = 0; z < nz; ++z) onSliceDone((int)z, out.getSlice(z));
}

PointCloud HipBackend::extractPointCloud(const Volume& vol, float threshold) {
    if (vol.empty()) return {};
    if (!m_ctx) check(CT_ERR_NODEVICE, "extractPointCloud");
    const int w = (int)vol.width, h = (int)vol.height, d = (int)vol.depth;
    const float* xc = vol.x_coords.size() >= vol.width ? vol.x_coords.data() : nullptr;
    const float* yc = vol.y_coords.size() >= vol.height ? vol.y_coords.data() : nullptr;
    const float* zc = vol.z_coords.size() >= vol.depth ? vol.z_coords.data() : nullptr;
    void* dv = nullptr;
    const size_t bytes = vol.data.size() * sizeof(float);
    check(ct_hip_alloc(m_ctx, bytes, &dv), "extractPointCloud");
    PointCloud cloud;
    try {
        check(ct_hip_upload(m_ctx, dv, vol.data.data(), bytes), "extractPointCloud");
        size_t n = 0;
        check(ct_hip_extract_point_cloud_dev(m_ctx, (const float*)dv, w, h, d, xc, yc, zc,
                                             threshold, nullptr, 0, &n),
              "extractPointCloud");
        cloud.resize(n);
        if (n)
            check(ct_hip_extract_point_cloud_dev(m_ctx, (const float*)dv, w, h, d, xc, yc, zc,
                                                 threshold, reinterpret_cast<float*>(cloud.data()),
                                                 n, &n),
                  "extractPointCloud");
    } catch (...) {
        ct_hip_free(m_ctx, dv);
        throw;
    }
    ct_hip_free(m_ctx, dv);
    return cloud;
}

HipBackend::Metrics HipBackend::volumeMetrics(const Volume& o, const Volume& r) {
    if (o.data.size() != r.data.size())
        throw std::invalid_argument("volumeMetrics: volumes differ in size");
    if (!m_ctx) check(CT_ERR_NODEVICE, "volumeMetrics");
    ct_metrics m{};
    check(ct_hip_volume_metrics(m_ctx, o.data.data(), r.data.data(), o.data.size(), &m),
          "volumeMetrics");
    Metrics out;
    out.mse = m.mse;
    out.mae = m.mae;
    out.psnr = m.psnr;
    out.max_abs = m.max_abs;
    out.vmin = m.vmin;
    out.vmax = m.vmax;
    out.count = (size_t)m.count;
    return out;
}

}  // namespace ct
