// IVolumeGenerator.h — the phantom generator interface, restated with the
// reference's signatures (QtCTReconstruction/IVolumeGenerator.h:16-27).
#pragma once

#include "CTTypes.h"
#include "Generator3D.h"

#include <string>

namespace ct {

class IVolumeGenerator {
public:
    virtual ~IVolumeGenerator() = default;
    virtual std::string name() const = 0;
    virtual bool isAvailable() const = 0;
    virtual Volume generateBrainHU(const Generator3D::Params& params) = 0;
};

}  // namespace ct
