// hip_backend_driver.cpp — drives HipBackend through the IReconstructionBackend
// interface obtained from BackendFactory, the way the reference's controller
// does (CtReconstructionController.cpp:1096-1146, 984-1003, 629-634), and
// writes every result as raw float32 files for the parity test. The file
// formats (FileIO.h) and the device metrics are driven the same way, and,
// given a seed and an ellipsoid count, the phantom through IVolumeGenerator
// as Generator3D::generateBrainHU calls it (Generator3D.cpp:41-44).
//
// usage: hip_backend_driver <volume.f32> <w> <h> <d> <num_angles> <filter> <threshold> <outdir>
//                           [<phantom seed> <ellipsoids>]
#include "FileIO.h"
#include "HipBackend.h"
#include "HipVolumeGenerator.h"

#include <memory>

#include <cstdio>
#include <cstdlib>
#include <fstream>
#include <iostream>
#include <string>
#include <vector>

using namespace ct;

static void dump(const std::string& path, const float* p, size_t n) {
    std::ofstream f(path, std::ios::binary);
    f.write(reinterpret_cast<const char*>(p), (std::streamsize)(n * sizeof(float)));
}

int main(int argc, char** argv) {
    if (argc != 9 && argc != 11) {
        std::cerr << "usage: " << argv[0] << " vol.f32 w h d na filter threshold outdir [seed n_ell]\n";
        return 2;
    }
    const size_t w = std::strtoul(argv[2], nullptr, 10), h = std::strtoul(argv[3], nullptr, 10),
                 d = std::strtoul(argv[4], nullptr, 10), na = std::strtoul(argv[5], nullptr, 10);
    const int filter = std::atoi(argv[6]);
    const float thr = std::strtof(argv[7], nullptr);
    const std::string out = argv[8];

    Volume vol(w, h, d);
    {
        std::ifstream f(argv[1], std::ios::binary);
        f.read(reinterpret_cast<char*>(vol.data.data()), (std::streamsize)(vol.data.size() * 4));
        if (!f) {
            std::cerr << "cannot read volume\n";
            return 3;
        }
    }
    auto backend = BackendFactory::createBestAvailable();
    if (!backend || !backend->isAvailable()) {
        std::cerr << "no available backend\n";
        return 4;
    }
    ReconstructionParams params;
    params.filter = static_cast<ReconstructionParams::FilterType>(filter);
    params.num_angles = na;
    try {
        Sinogram s = backend->computeSinogram(vol.getSlice(0), na, w);
        dump(out + "/sino0.f32", s.data.data.data(), s.data.data.size());
        Buffer2D r = backend->reconstructSlice(s, w, params);
        dump(out + "/slice0.f32", r.data.data(), r.data.size());

        Volume rec;
        int calls = 0;
        backend->reconstructVolume(vol, rec, params, [&](int, const Buffer2D&) { ++calls; });
        dump(out + "/volume.f32", rec.data.data(), rec.data.size());

        // Consumer path: sinograms of every slice, then FBP from them
        Volume sinos(w, na, d);
        for (size_t z = 0; z < d; ++z) {
            Sinogram sz = backend->computeSinogram(vol.getSlice(z), na, w);
            std::copy(sz.data.data.begin(), sz.data.data.end(), sinos.data.begin() + z * na * w);
        }
        Volume rec2;
        backend->reconstructVolumeFromSinograms(sinos, s.angles_deg, rec2, params, nullptr);
        dump(out + "/fromsino.f32", rec2.data.data(), rec2.data.size());
        // the interface's default body (one reconstructSlice per z), which a
        // backend without its own override would run
        Volume rec_default;
        int default_calls = 0;
        backend->IReconstructionBackend::reconstructVolumeFromSinograms(
            sinos, s.angles_deg, rec_default, params, [&](int, const Buffer2D&) { ++default_calls; });
        dump(out + "/fromsino_default.f32", rec_default.data.data(), rec_default.data.size());

        vol.x_coords.resize(w);
        for (size_t x = 0; x < w; ++x) vol.x_coords[x] = 0.5f * (float)x;
        PointCloud pc = backend->extractPointCloud(vol, thr);
        dump(out + "/points.f32", reinterpret_cast<const float*>(pc.data()), pc.size() * 4);

        // files: the sinogram stack exported and loaded back, then the
        // consumer path from the loaded stack; volume and point cloud NPYs
        HipBackend* hb = dynamic_cast<HipBackend*>(backend.get());
        if (!hb) throw std::runtime_error("best available backend is not HipBackend");
        if (!SinogramFiles::exportAll(*hb, vol, na, out + "/sinos.npy"))
            throw std::runtime_error("exportAll failed");
        LoadedSinograms ls;
        if (!SinogramFiles::load(out + "/sinos.npy", ls)) throw std::runtime_error("load failed");
        Volume rec3;
        backend->reconstructVolumeFromSinograms(ls.sinograms, ls.angles_deg, rec3, params, nullptr);
        dump(out + "/fromfile.f32", rec3.data.data(), rec3.data.size());
        Volume back;
        const bool vol_rt = FileIO::saveVolumeNPY(rec, out + "/volume.npy") &&
                            FileIO::loadVolumeNPY(out + "/volume.npy", back, false) &&
                            back.data == rec.data && back.x_coords.size() == w;
        PointCloud pc_back;
        const bool pc_rt = pc.empty() || (FileIO::savePointCloudNPY(pc, out + "/points.npy") &&
                                          FileIO::loadPointCloudNPY(out + "/points.npy", pc_back) &&
                                          pc_back.size() == pc.size());
        const HipBackend::Metrics m = hb->volumeMetrics(vol, rec);

        // empty inputs give empty results, no exception
        Sinogram e = backend->computeSinogram(Buffer2D(), na, w);
        Buffer2D er = backend->reconstructSlice(Sinogram(), w, params);

        std::ofstream meta(out + "/meta.txt");
        meta << backend->name() << "\n"
             << calls << "\n"
             << pc.size() << "\n"
             << (e.data.empty() && er.empty()) << "\n"
             << backend->lastSinogramTimeMs() << "\n"
             << backend->lastReconstructionTimeMs() << "\n"
             << vol_rt << "\n"
             << pc_rt << "\n";
        meta.precision(17);
        meta << default_calls << "\n";
        meta << m.mse << "\n" << m.mae << "\n" << m.psnr << "\n" << m.max_abs << "\n" << m.vmin
             << "\n" << m.vmax << "\n" << m.count << "\n";
        if (argc == 11) {
            std::unique_ptr<IVolumeGenerator> gen = std::make_unique<HipVolumeGenerator>();
            Generator3D::Params gp;
            gp.shape = {w, h, d};
            gp.seed = (uint32_t)std::strtoul(argv[9], nullptr, 10);
            gp.num_ellipsoids = std::strtoul(argv[10], nullptr, 10);
            Volume ph = gen->generateBrainHU(gp);
            dump(out + "/phantom.f32", ph.data.data(), ph.data.size());
            meta << gen->name() << "\n"
                 << gen->isAvailable() << "\n"
                 << ph.width << " " << ph.height << " " << ph.depth << "\n"
                 << ph.x_coords.front() << " " << ph.x_coords.back() << " " << ph.z_coords.size()
                 << "\n";
        }
    } catch (const std::exception& ex) {
        std::cerr << "exception: " << ex.what() << "\n";
        return 5;
    }
    return 0;
}
