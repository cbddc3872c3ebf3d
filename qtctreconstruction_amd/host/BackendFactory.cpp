// BackendFactory.cpp — backend selection with the HIP backend registered.
// Mirrors BackendFactory::create / createBestAvailable
// (QtCTReconstruction/BackendFactory.cpp:62-92):
//   * HIP is a new BackendType after the reference's four;
//   * CUDA maps to HIP in a build without CUDA, because the Consumer UI
//     forces backend 2 (main_consumer.qml:22-26);
//   * createBestAvailable tries HIP first.
// Serial / OpenMP / Hybrid live in the reference tree; in this library they
// return nullptr (the controller then falls back, CtReconstructionController.cpp:1098-1101).
#include "HipBackend.h"

namespace ct {

// The interface's default consumer path (BackendFactory.cpp:17-60 of the
// reference): sinograms are (depth = nz, height = na, width = bins); every z
// is reconstructed by reconstructSlice onto bins x bins; the Sinogram it is
// handed carries the slice's own min/max as original_min/max_hu and spacing 1.
void IReconstructionBackend::reconstructVolumeFromSinograms(
    const Volume& sinograms, const std::vector<float>& angles_deg, Volume& out_reconstruction,
    const ReconstructionParams& params, std::function<void(int, const Buffer2D&)> onSliceDone) {
    const size_t nz = sinograms.depth, na = sinograms.height, bins = sinograms.width;
    if (nz == 0 || na == 0 || bins == 0) return;
    out_reconstruction.assign(bins, bins, nz, 0.0f);
    for (size_t z = 0; z < nz; ++z) {
        Sinogram s;
        s.data = sinograms.getSlice(z);
        s.angles_deg = angles_deg;
        s.detector_spacing_mm = 1.0f;
        float lo = s.data.data.empty() ? 0.0f : s.data.data[0];
        float hi = lo;
        for (const float v : s.data.data) {
            lo = v < lo ? v : lo;
            hi = v > hi ? v : hi;
        }
        s.original_min_hu = lo;
        s.original_max_hu = hi;
        const Buffer2D recon = reconstructSlice(s, bins, params);
        out_reconstruction.setSlice(z, recon);
        if (onSliceDone) onSliceDone(static_cast<int>(z), recon);
    }
}

std::shared_ptr<IReconstructionBackend> BackendFactory::create(BackendType type) {
    if (type == BackendType::HIP || type == BackendType::CUDA) {
        auto b = std::make_shared<HipBackend>();
        return b;
    }
    return nullptr;
}

std::shared_ptr<IReconstructionBackend> BackendFactory::createBestAvailable() {
    auto hip = create(BackendType::HIP);
    if (hip && hip->isAvailable()) return hip;
    return create(BackendType::OpenMP);
}

}  // namespace ct
