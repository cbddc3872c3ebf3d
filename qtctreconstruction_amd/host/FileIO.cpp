// FileIO.cpp — FileIO and SinogramFiles over the C ABI (see FileIO.h).
#include "FileIO.h"

#include "HipBackend.h"

#include "../../include/ct_hip.h"

#include <stdexcept>

namespace ct {

namespace {
std::vector<float> linspace(float start, float stop, size_t count) {
    if (count == 0) return {};
    if (count == 1) return {start};
    std::vector<float> v(count);
    const float step = (stop - start) / static_cast<float>(count - 1);
    for (size_t i = 0; i < count; ++i) v[i] = start + step * static_cast<float>(i);
    return v;
}
}  // namespace

bool FileIO::saveVolumeNPY(const Volume& v, const std::string& path) {
    if (v.width == 0 || v.height == 0 || v.depth == 0) return false;
    return ct_hip_save_volume_npy(v.data.data(), v.width, v.height, v.depth, path.c_str()) == CT_OK;
}

bool FileIO::loadVolumeNPY(const std::string& path, Volume& out, bool xyz) {
    size_t dims[3];
    if (ct_hip_load_volume_npy(path.c_str(), xyz ? 1 : 0, nullptr, 0, dims) != CT_OK) return false;
    Volume v;
    v.assign(dims[0], dims[1], dims[2], 0.0f);
    if (ct_hip_load_volume_npy(path.c_str(), xyz ? 1 : 0, v.data.data(), v.data.size(), dims) != CT_OK)
        return false;
    v.x_coords = linspace(-100.0f, 100.0f, v.width);
    v.y_coords = linspace(-100.0f, 100.0f, v.height);
    v.z_coords = linspace(-100.0f, 100.0f, v.depth);
    out = std::move(v);
    return true;
}

bool FileIO::savePointCloudNPY(const PointCloud& cloud, const std::string& path) {
    if (cloud.empty()) return false;
    return ct_hip_save_point_cloud_npy(reinterpret_cast<const float*>(cloud.data()), cloud.size(),
                                       path.c_str()) == CT_OK;
}

bool FileIO::loadPointCloudNPY(const std::string& path, PointCloud& out) {
    size_t n = 0;
    if (ct_hip_load_point_cloud_npy(path.c_str(), nullptr, 0, &n) != CT_OK) return false;
    PointCloud pc(n);
    if (ct_hip_load_point_cloud_npy(path.c_str(), reinterpret_cast<float*>(pc.data()), n, &n) != CT_OK)
        return false;
    out = std::move(pc);
    return true;
}

bool SinogramFiles::exportAll(HipBackend& backend, const Volume& v, size_t na,
                              const std::string& path) {
    if (v.empty() || na == 0) return false;
    const int rc = ct_hip_export_sinograms(backend.context(), v.data.data(), (int)v.width,
                                           (int)v.height, (int)v.depth, (int)na, path.c_str());
    if (rc == CT_ERR_IO) return false;
    if (rc != CT_OK)
        throw std::runtime_error("exportAllSinograms failed (" + std::to_string(rc) + "): " +
                                 ct_hip_last_error(backend.context()));
    return true;
}

bool SinogramFiles::load(const std::string& path, LoadedSinograms& out) {
    ct_sino_info info{};
    if (ct_hip_load_sinograms(path.c_str(), nullptr, 0, nullptr, &info) != CT_OK) return false;
    LoadedSinograms r;
    r.sinograms.assign(info.bins, info.na, info.nz, 0.0f);
    r.angles_deg.resize(info.na);
    if (ct_hip_load_sinograms(path.c_str(), r.sinograms.data.data(), r.sinograms.data.size(),
                              r.angles_deg.data(), &info) != CT_OK)
        return false;
    r.detector_spacing_mm = info.detector_spacing_mm;
    r.min_hu = info.min_hu;
    r.max_hu = info.max_hu;
    out = std::move(r);
    return true;
}

}  // namespace ct
