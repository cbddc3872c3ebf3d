// IReconstructionBackend.h — the backend plugin interface, restated with the
// reference's signatures (QtCTReconstruction/IReconstructionBackend.h:10-64) so
// HipBackend is a drop-in next to SerialBackend / OpenMPBackend / CUDABackend.
// BackendType gains HIP after the reference's four values, which the UI
// stores as integers (main.qml:86-92), so existing values keep their meaning.
#pragma once

#include "CTTypes.h"

#include <functional>
#include <memory>
#include <string>

namespace ct {

class IReconstructionBackend {
public:
    virtual ~IReconstructionBackend() = default;
    virtual std::string name() const = 0;
    virtual bool isAvailable() const = 0;
    virtual Sinogram computeSinogram(const Buffer2D& slice, size_t num_angles,
                                     size_t detector_bins, bool use_parallel = true) = 0;
    virtual Buffer2D reconstructSlice(const Sinogram& sinogram, size_t output_size,
                                      const ReconstructionParams& params) = 0;
    virtual void reconstructVolume(
        const Volume& input_volume, Volume& out_reconstruction, const ReconstructionParams& params,
        std::function<void(int slice_idx, const Buffer2D& recon_slice)> onSliceDone) = 0;
    // Default body as the reference's (BackendFactory.cpp:17-60): one
    // reconstructSlice per z onto bins x bins, each slice's min/max taken from
    // its sinogram; defined in BackendFactory.cpp. HipBackend overrides it with
    // one batched device call.
    virtual void reconstructVolumeFromSinograms(
        const Volume& sinograms, const std::vector<float>& angles_deg, Volume& out_reconstruction,
        const ReconstructionParams& params,
        std::function<void(int slice_idx, const Buffer2D& recon_slice)> onSliceDone);
    virtual PointCloud extractPointCloud(const Volume& vol, float threshold) = 0;
    virtual double lastSinogramTimeMs() const = 0;
    virtual double lastReconstructionTimeMs() const { return 0.0; }
};

class BackendFactory {
public:
    enum class BackendType { Serial, OpenMP, CUDA, Hybrid, HIP };
    static std::shared_ptr<IReconstructionBackend> create(BackendType type);
    static std::shared_ptr<IReconstructionBackend> createBestAvailable();
};

}  // namespace ct
