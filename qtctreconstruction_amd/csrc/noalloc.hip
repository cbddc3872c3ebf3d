// noalloc.hip — reconstructVolumeFromSinograms on device memory without any
// device allocation, for callers that cannot allocate (code under HIP graph
// capture, tight serving loops, harnesses that forbid hipMalloc).
//
// Its only scratch (the angle table, the filter taps and FFT twiddles of
// either row filter form, and the filtered rows) lives in one file-scope
// array of kScratchBytes per GPU,
// which the device holds from the library's load to the process's end. The
// filtered rows go through it in slabs of whole 4-slice z-blocks, so any
// depth fits; a shape whose single z-block of filtered rows (na x sq x 16
// bytes) does not fit is refused. The kernels, and the z-blocks the FFT pairs
// its rows in, are those of ct_hip_reconstruct_from_sinograms_dev, so the
// output is the same bit for bit.
//
// Every call on one GPU shares that GPU's array and leaves the last call's
// contents in it: calls on one context are serialised by its lock, calls on
// two contexts of one GPU that overlap race on the array.
#include "ctx.h"

#include <cmath>
#include <cstring>
#include <mutex>

using namespace cthip;

namespace cthost {

// 64 MiB: the filtered rows of one z-block take na x sq x 16 bytes (8.4 MB
// at 720 angles x 725 padded bins), so 7 z-blocks per slab at that shape.
constexpr size_t kScratchBytes = size_t(64) << 20;
// Head of the array: the angle table (2 x kNoAllocMaxAngles floats), the
// LDS form's filter taps (kMaxFFT / 2 + 1 floats) and twiddles (kMaxFFT - 1
// float2), the wave form's twiddles and filter (kWaveTabEntries float2, then
// kWaveHsFloats floats), each 256-byte aligned; the filtered rows take the
// rest.
constexpr int kNoAllocMaxAngles = 8192;
constexpr size_t kTrigOff = 0;
constexpr size_t kTapsOff = kTrigOff + 2 * sizeof(float) * kNoAllocMaxAngles;
constexpr size_t kTwOff = kTapsOff + (sizeof(float) * (kMaxFFT / 2 + 1) + 255) / 256 * 256;
constexpr size_t kWaveOff = kTwOff + (sizeof(float2) * (kMaxFFT - 1) + 255) / 256 * 256;
constexpr size_t kWaveBytes = sizeof(float2) * kWaveTabEntries + sizeof(float) * kWaveHsFloats;
constexpr size_t kFiltOff = kWaveOff + (kWaveBytes + 255) / 256 * 256;
static_assert(kFiltOff < kScratchBytes, "scratch head larger than the scratch");

static __device__ __align__(256) unsigned char scratch[kScratchBytes];

// The array's device address on each GPU (its own copy per device), taken
// once per device with that device current.
constexpr int kMaxDevices = 64;
static std::mutex g_scratch_mu;
static void* g_scratch_ptr[kMaxDevices] = {};
static hipError_t g_scratch_err[kMaxDevices] = {};
static bool g_scratch_done[kMaxDevices] = {};

// Called by ct_hip_create with `device` current.
void noalloc_init(int device) {
    if (device < 0 || device >= kMaxDevices) return;
    std::lock_guard<std::mutex> lk(g_scratch_mu);
    if (g_scratch_done[device]) return;
    void* p = nullptr;
    const hipError_t e = hipGetSymbolAddress(&p, HIP_SYMBOL(scratch));
    g_scratch_err[device] = e;
    g_scratch_ptr[device] = e == hipSuccess ? p : nullptr;
    if (e != hipSuccess) (void)hipGetLastError();
    g_scratch_done[device] = true;
}

static char* scratch_of(int device) {
    std::lock_guard<std::mutex> lk(g_scratch_mu);
    if (device < 0 || device >= kMaxDevices || !g_scratch_done[device])
        throw StatusError(CT_ERR_HIP, "noalloc: scratch address not taken for this device");
    if (!g_scratch_ptr[device])
        throw StatusError(CT_ERR_HIP, std::string("noalloc: hipGetSymbolAddress failed: ") +
                                          hipGetErrorString(g_scratch_err[device]));
    return static_cast<char*>(g_scratch_ptr[device]);
}

// The z-blocks of filtered rows one slab of the scratch holds (0: none).
static size_t noalloc_zblocks(int na, int sq) {
    return (kScratchBytes - kFiltOff) / ((size_t)na * sq * sizeof(float4));
}

static void from_sinos_noalloc(ct_hip_ctx* c, const float* d_sinos, int nz, int na, int bins,
                               const float* angles_deg, int filter, float* d_out) {
    need(d_sinos && angles_deg && d_out, "reconstructVolumeFromSinograms: null pointer");
    need(nz > 0 && na > 0 && bins > 0, "reconstructVolumeFromSinograms: empty input");
    check_fbp(c, bins);
    if (c->filter_backend != CT_FILTER_BACKEND_FFT_F32)
        throw StatusError(CT_ERR_UNSUPPORTED,
                          "the allocation-free form runs the FFT row filter only (the Toeplitz "
                          "GEMM backends keep device tables of their own)");
    if (na > kNoAllocMaxAngles)
        throw StatusError(CT_ERR_UNSUPPORTED, "the allocation-free form takes at most " +
                                                  std::to_string(kNoAllocMaxAngles) + " angles");
    const Geom g = geometry(bins);
    const size_t zbs = noalloc_zblocks(na, g.sq);
    if (zbs == 0)
        throw StatusError(CT_ERR_UNSUPPORTED,
                          "one z-block of filtered rows (" + std::to_string(na) + " angles x " +
                              std::to_string(g.sq) + " bins x 16 B) exceeds the " +
                              std::to_string(kScratchBytes >> 20) + " MiB scratch");
    char* base = scratch_of(c->device);
    float* d_cos = reinterpret_cast<float*>(base + kTrigOff);
    float* d_sin = d_cos + na;
    float* d_H = reinterpret_cast<float*>(base + kTapsOff);
    float2* d_tw = reinterpret_cast<float2*>(base + kTwOff);
    float4* d_filt = reinterpret_cast<float4*>(base + kFiltOff);
    float2* d_wtw = reinterpret_cast<float2*>(base + kWaveOff);
    float* d_whs = reinterpret_cast<float*>(d_wtw + kWaveTabEntries);
    // Serial's trig (SerialBackend.cpp:198-204) and the row filter's tables
    // (taps with the Nyquist bin dropped, FFT.cpp:87-120), made on the host
    // as the allocating entry point makes them and copied in; the stream is
    // drained before return, so these host arrays outlive the copies.
    static thread_local float trig[2 * kNoAllocMaxAngles];
    static thread_local float taps[kMaxFFT / 2 + 1];
    static thread_local float2 wtw[kWaveTabEntries];
    static thread_local float whs[kWaveHsFloats];
    for (int a = 0; a < na; ++a) {
        const float th = deg2rad(angles_deg[a]);
        trig[a] = std::cos(th);
        trig[na + a] = std::sin(th);
    }
    hip_check(hipMemcpyAsync(d_cos, trig, sizeof(float) * 2 * na, hipMemcpyHostToDevice, c->stream),
              "noalloc: angle table");
    const int wave_r = wave_filter_r(bins, g.sq, g.pad, g.nfft);
    FiltArgs fa;
    if (wave_r > 0) {
        wave_tables_host(g, filter, wtw, whs);
        hip_check(hipMemcpyAsync(d_wtw, wtw, sizeof(wtw), hipMemcpyHostToDevice, c->stream),
                  "noalloc: wave twiddles");
        hip_check(hipMemcpyAsync(d_whs, whs, sizeof(whs), hipMemcpyHostToDevice, c->stream),
                  "noalloc: wave filter");
        fa = make_filt_args(d_sinos, nz, na, bins, g.sq, g.pad, g.nfft, d_whs, d_wtw, d_filt);
        fa.wave_r = wave_r;
    } else {
        const int L = lds_filter_length(g);
        lds_taps_host(g, filter, L, taps);
        hip_check(hipMemcpyAsync(d_H, taps, sizeof(float) * (L / 2 + 1), hipMemcpyHostToDevice,
                                 c->stream),
                  "noalloc: filter taps");
        hip_check(launch_twiddles(L, d_tw, c->stream), "noalloc: twiddles");
        fa = make_filt_args(d_sinos, nz, na, bins, g.sq, g.pad, L, d_H, d_tw, d_filt);
    }
    // the tuned configuration when this (N, na) class has been tuned, else the
    // context's: every configuration gives the same bits, and tuning allocates
    auto it = c->tuned.find({bins, na});
    const int cfg = c->bp_config_fixed ? c->bp_config
                                       : (it != c->tuned.end() ? it->second : c->bp_config);
    const float scale = 3.14159265358979323846f / (2.0f * (float)na);  // SerialBackend.cpp:246
    const int slab = (int)std::min<size_t>(zbs * kZB, (size_t)nz);
    for (int z0 = 0; z0 < nz; z0 += slab) {
        const int n = std::min(slab, nz - z0);
        FiltArgs slab = fa;
        slab.sino = d_sinos + (size_t)z0 * na * bins;
        slab.nz = n;
        hip_check(launch_filter(slab, c->stream), "noalloc: row filter");
        hip_check(launch_backproject(d_filt, n, na, g.sq, bins, d_cos, d_sin, scale, nullptr,
                                     d_out + (size_t)z0 * bins * bins, c->stream, cfg),
                  "noalloc: backprojection");
    }
    hip_check(hipStreamSynchronize(c->stream), "sync");
}

}  // namespace cthost

using namespace cthost;

extern "C" int ct_hip_reconstruct_from_sinograms_dev_noalloc(ct_hip_ctx* c, const float* d_sinos,
                                                             int nz, int na, int bins,
                                                             const float* angles_deg, int filter,
                                                             float* d_out) {
    return guarded(c, [&] { from_sinos_noalloc(c, d_sinos, nz, na, bins, angles_deg, filter, d_out); });
}
