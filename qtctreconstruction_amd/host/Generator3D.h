// Generator3D.h — the phantom generator's parameter types, restated with the
// reference's names, fields and defaults (QtCTReconstruction/Generator3D.h:12-44)
// so that HipVolumeGenerator implements IVolumeGenerator with the reference's
// signature. Only the types are restated: the reference's Generator3D picks
// its implementation in setBackend (Generator3D.cpp:18-39), where a HIP
// branch constructs HipVolumeGenerator.
#pragma once

#include "CTTypes.h"

#include <array>
#include <cstddef>
#include <cstdint>

namespace ct {

class Generator3D {
public:
    enum class BackendType { CPU, CUDA };

    struct TissueParams {
        float hu_min = -20.0f;
        float hu_max = 45.0f;
        float size_min_mm = 3.0f;
        float size_max_mm = 10.0f;
        bool peripheral_only = false;
        float peripheral_radius_min = 0.0f;  // fraction of brain_radius_mm
        float peripheral_radius_max = 0.0f;  // fraction of brain_radius_mm
        float probability = 1.0f;
    };

    struct Params {
        std::array<size_t, 3> shape = {128, 128, 64};  // [x, y, z]
        float brain_radius_mm = 50.0f;
        size_t num_ellipsoids = 120;
        uint32_t seed = 42;
        TissueParams soft_tissue{};
        TissueParams bone_tissue{200.0f, 1000.0f, 2.0f, 8.0f, true, 0.7f, 0.9f, 0.15f};
    };
};

}  // namespace ct
