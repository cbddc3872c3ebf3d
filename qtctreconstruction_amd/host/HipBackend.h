// HipBackend.h — IReconstructionBackend on the MI355X library (libct_hip.so).
// A sibling of the reference's CUDABackend (CUDABackend.h:10-40): same
// virtuals, same ownership (the backend owns its device state and caches it
// across calls), same threading contract (calls may come from several
// threads; the C-ABI context serialises them). Numerics follow SerialBackend.
#pragma once

#include "IReconstructionBackend.h"

struct ct_hip_ctx;

namespace ct {

class HipBackend : public IReconstructionBackend {
public:
    explicit HipBackend(int device = 0, size_t hbm_budget_bytes = 0);
    ~HipBackend() override;
    HipBackend(const HipBackend&) = delete;
    HipBackend& operator=(const HipBackend&) = delete;

    std::string name() const override { return "HIP (MI355X)"; }
    bool isAvailable() const override;

    Sinogram computeSinogram(const Buffer2D& slice, size_t num_angles, size_t detector_bins,
                             bool use_parallel = true) override;
    Buffer2D reconstructSlice(const Sinogram& sinogram, size_t output_size,
                              const ReconstructionParams& params) override;
    void reconstructVolume(const Volume& input_volume, Volume& out_reconstruction,
                           const ReconstructionParams& params,
                           std::function<void(int, const Buffer2D&)> onSliceDone) override;
    void reconstructVolumeFromSinograms(const Volume& sinograms,
                                        const std::vector<float>& angles_deg,
                                        Volume& out_reconstruction,
                                        const ReconstructionParams& params,
                                        std::function<void(int, const Buffer2D&)> onSliceDone) override;
    PointCloud extractPointCloud(const Volume& vol, float threshold) override;
    double lastSinogramTimeMs() const override { return m_lastSinogramTimeMs; }
    double lastReconstructionTimeMs() const override { return m_lastReconstructionTimeMs; }

    // Not virtuals of the interface: the controller's reconstruction metrics
    // (CtReconstructionController.cpp:1199-1277) computed on the device, and
    // the C-ABI context for the file helpers of FileIO.h.
    struct Metrics {
        double mse = 0, mae = 0, psnr = 0;
        float max_abs = 0, vmin = 0, vmax = 0;
        size_t count = 0;
    };
    Metrics volumeMetrics(const Volume& original, const Volume& reconstruction);
    ct_hip_ctx* context() const { return m_ctx; }

private:
    void check(int rc, const char* what) const;  // throws std::runtime_error, as GPUCHECK does
    ct_hip_ctx* m_ctx = nullptr;
    double m_lastSinogramTimeMs = 0.0;
    double m_lastReconstructionTimeMs = 0.0;
};

}  // namespace ct
