// HipVolumeGenerator.h — IVolumeGenerator on the MI355X library: a sibling
// of the reference's Generator3DGPU (Generator3DGPU.h:9-17, .cu:124-177).
//
// It rasterises the project phantom of the benchmark configurations (additive
// ellipsoids with ZYZ rotations, seeded by Params::seed, Params::num_ellipsoids
// of them, on a Params::shape grid) in one device launch and copies the
// volume back once; the reference's GPU generator runs one launch and one
// synchronous copy per slice. The brain-HU tissue parameters (HU ranges,
// sizes, the bone probability and brain_radius_mm) describe the reference's
// own phantom, which the configurations do not use: they do not change the
// voxels. brain_radius_mm sets the coordinate axes, as in the reference
// (linspace(-R, R, n) per axis).
#pragma once

#include "IVolumeGenerator.h"

struct ct_hip_ctx;

namespace ct {

class HipVolumeGenerator : public IVolumeGenerator {
public:
    explicit HipVolumeGenerator(int device = 0);
    ~HipVolumeGenerator() override;
    HipVolumeGenerator(const HipVolumeGenerator&) = delete;
    HipVolumeGenerator& operator=(const HipVolumeGenerator&) = delete;

    std::string name() const override { return "GPU (HIP, MI355X)"; }
    bool isAvailable() const override;
    Volume generateBrainHU(const Generator3D::Params& params) override;

private:
    ct_hip_ctx* m_ctx = nullptr;
};

}  // namespace ct
