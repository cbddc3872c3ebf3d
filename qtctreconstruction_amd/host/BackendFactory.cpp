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
