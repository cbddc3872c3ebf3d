// ctx.h — the context behind the C ABI (private to the library): device
// allocations with HBM-budget accounting, cached device state and the helper
// functions shared by capi.cpp, batch.cpp and serve.cpp.
#pragma once

#include "../../include/ct_hip.h"
#include "ct_internal.h"

#include <cstring>
#include <map>
#include <mutex>
#include <stdexcept>
#include <string>
#include <tuple>
#include <unordered_map>
#include <vector>

namespace cthost {

struct Buf {
    void* p = nullptr;
    size_t cap = 0;
};

struct StatusError : std::runtime_error {
    int code;
    StatusError(int c, const std::string& m) : std::runtime_error(m), code(c) {}
};

struct Geom {
    int bins, sq, pad, nfft;
};

struct BatchGraph;  // batch.cpp

// Largest number of volumes in one batch.
constexpr int kMaxBatch = 256;
// Distinct FFT lengths in one batch (64 .. 4096).
constexpr int kMaxFftClasses = 8;

// Launch state of one stream that runs batches (batch.cpp).
struct BatchSlot {
    bool ready = false, own_stream = false;
    hipStream_t stream = nullptr;
    hipEvent_t ev_start = nullptr, ev_done = nullptr;
    cthip::FiltJob* ftab = nullptr;  // device work lists (kMaxBatch entries)
    cthip::BpJob* btab = nullptr;
    int* hdr = nullptr;
    int* counters = nullptr;
    cthip::FiltJob* h_ftab = nullptr;  // pinned staging
    cthip::BpJob* h_btab = nullptr;
    int* h_hdr = nullptr;
    std::map<std::vector<int>, hipGraphExec_t> graphs;  // keyed by the launch shape
    ~BatchSlot();
};

}  // namespace cthost

using cthost::Buf;

struct ct_hip_ctx {
    int device = 0;
    hipStream_t stream = nullptr;
    size_t budget = 0;  // 0 = unlimited
    size_t in_use = 0;
    std::recursive_mutex mu;
    std::string err;
    std::unordered_map<void*, size_t> sizes;  // every live device allocation
    // workspaces (grow-only, reused across calls)
    Buf vol, row, col, side, sino, filt, out, keys, mm, ell, pc_cnt, pc_pts, pc_coord, met;
    // cached device state
    std::map<std::pair<int, int>, float*> filters;  // (nfft, type) -> H[nfft/2+1]
    std::map<int, float2*> twiddles;                // nfft -> tw[nfft-1], stage order
    struct Trig {
        float* cos = nullptr;
        float* sin = nullptr;
        // angle indices for the Radon stage: n_row angles with |sin| <= |cos|
        // (row walk over the image), then the rest (over the transpose)
        int* alist = nullptr;
        int n_row = 0;
        uint64_t last_use = 0;  // trig_tick of the last call that used the table
    };
    std::map<std::vector<uint32_t>, Trig> trig;  // keyed by the angle list's bit pattern
    Trig cur;                                     // table of the current plain call
    uint64_t trig_tick = 0;
    hipEvent_t ev[4] = {nullptr, nullptr, nullptr, nullptr};
    double timings[4] = {0, 0, 0, 0};
    // Radon stage breakdown: events before the walk over the image, between
    // the two walks and after the walk over the transpose
    hipEvent_t evr[3] = {nullptr, nullptr, nullptr};
    double radon_ms[2] = {0, 0};
    bool radon_timed = false;
    int filter_backend = 0;  // CT_FILTER_BACKEND_*
    int bp_config = 0;       // backprojection tile configuration (kBpConfigs index)
    bool bp_config_fixed = false;  // set by ct_hip_set_bp_config; else tuned per (N, na)
    struct Toep {
        int Np = 0, Kp = 0;
        unsigned short *hi = nullptr, *lo = nullptr;
        unsigned char* f8 = nullptr;
        float* tmax = nullptr;
    };
    std::map<std::tuple<int, int, int>, Toep> toeplitz;  // (nfft, type, bins)
    Buf rmax;
    // batched form (batch.cpp)
    std::map<std::pair<int, int>, int> tuned;  // (N, na) -> backprojection config
    bool tune_cache_loaded = false;            // CT_HIP_TUNE_CACHE read into `tuned`
    cthost::BatchGraph* batch = nullptr;
};

namespace cthost {

void hip_check(hipError_t e, const char* what);
void* dmalloc(ct_hip_ctx* c, size_t bytes);
void dfree(ct_hip_ctx* c, void* p);
void release(ct_hip_ctx* c, Buf& b);
void need(bool ok, const char* msg);
int norm_filter(int f);
Geom geometry(int bins);
void check_fft(const Geom& g);
void check_fbp(const ct_hip_ctx* c, int bins);
void trim_trig(ct_hip_ctx* c);
ct_hip_ctx::Trig get_trig(ct_hip_ctx* c, const float* angles_deg, int na);
float* get_filter(ct_hip_ctx* c, int nfft, int type);
float2* get_twiddles(ct_hip_ctx* c, int nfft);
void bind_device(ct_hip_ctx* c);
void batch_release(ct_hip_ctx* c);
int autotune(ct_hip_ctx* c, int N, int na, std::vector<double>* times);
void slot_init(ct_hip_ctx* c, BatchSlot* s, bool own_stream);
void slot_release(ct_hip_ctx* c, BatchSlot* s);
size_t job_filt_bytes(const ct_fbp_job& j);
void enqueue_batch(ct_hip_ctx* c, BatchSlot* s, const ct_fbp_job* jobs, int nj,
                   float4* const* filts, int flags);
void write_sinogram_files(const std::string& npy_path, const float* sinos, int nz, int na,
                          int bins, float min_hu, float max_hu);

template <typename T>
T* ensure(ct_hip_ctx* c, Buf& b, size_t bytes) {
    if (b.cap < bytes) {
        if (b.p) {
            hip_check(hipStreamSynchronize(c->stream), "sync before realloc");
            dfree(c, b.p);
            b.p = nullptr;
            b.cap = 0;
        }
        b.p = dmalloc(c, bytes);
        b.cap = bytes;
    }
    return static_cast<T*>(b.p);
}

template <typename F>
int guarded(ct_hip_ctx* c, F&& f) {
    if (!c) return CT_ERR_INVALID;
    std::lock_guard<std::recursive_mutex> lk(c->mu);
    try {
        bind_device(c);
        f();
        c->err.clear();
        return CT_OK;
    } catch (const StatusError& e) {
        c->err = e.what();
        return e.code;
    } catch (const std::bad_alloc&) {
        c->err = "host allocation failed";
        return CT_ERR_NOMEM;
    } catch (const std::exception& e) {
        c->err = e.what();
        return CT_ERR_INVALID;
    }
}

}  // namespace cthost
