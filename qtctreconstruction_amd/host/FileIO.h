// FileIO.h — the reference's file formats for C++ callers, over the C ABI of
// libct_hip.so (host code; no device needed).
//   FileIO           NPY volumes and point clouds, the static class of
//                    FileIO.h:9-21 with the same signatures and bool results
//                    (slice PNG/BMP export belongs to the UI and is not here)
//   SinogramFiles    the controller's sinogram stack on disk: NPY (nz,na,bins)
//                    plus the .meta.json sidecar
//                    (CtReconstructionController.cpp:1353-1482, 1494-1616)
#pragma once

#include "CTTypes.h"

#include <string>
#include <vector>

namespace ct {

class FileIO {
public:
    static bool saveVolumeNPY(const Volume& volume, const std::string& path_utf8);
    // Axis coordinates are set to linspace(-100, 100, n), as FileIO.cpp:141-143.
    static bool loadVolumeNPY(const std::string& path_utf8, Volume& out_volume,
                              bool input_is_xyz_layout);
    static bool savePointCloudNPY(const PointCloud& cloud, const std::string& path_utf8);
    static bool loadPointCloudNPY(const std::string& path_utf8, PointCloud& out_cloud);
};

class HipBackend;

struct LoadedSinograms {
    Volume sinograms;  // width = bins, height = na, depth = nz
    std::vector<float> angles_deg;
    float detector_spacing_mm = 1.0f;
    float min_hu = 0.0f, max_hu = 0.0f;
};

struct SinogramFiles {
    // exportAllSinograms for a volume in memory: computeSinogram of every
    // slice (bins = width) on the backend's device, then the two files.
    // Throws std::runtime_error on a device failure; false when a file
    // cannot be written.
    static bool exportAll(HipBackend& backend, const Volume& volume, size_t num_angles,
                          const std::string& npy_path);
    // loadSinograms without the dialog: false when the NPY is missing or bad.
    static bool load(const std::string& npy_path, LoadedSinograms& out);
};

}  // namespace ct
