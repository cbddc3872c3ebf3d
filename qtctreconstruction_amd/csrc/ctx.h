// ctx.h — the context behind the C ABI (private to the library): device
// allocations with HBM-budget accounting, cached device state and the helper
// functions shared by capi.cpp, batch.cpp and serve.cpp.
#pragma once

#include "../../include/ct_hip.h"
#include "ct_internal.h"

#include <atomic>
#include <condition_variable>
#include <cstring>
#include <map>
#include <memory>
#include <mutex>
#include <thread>
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

// Parallel memcpy between pageable caller memory and pinned staging: the
// calling thread and `workers` pool threads each take chunks of one copy at a
// time (hostpath.cpp); one copy runs at a time, whichever thread asks. Below
// 1 MB the caller copies alone.
class CopyPool {
public:
    explicit CopyPool(int workers);
    ~CopyPool();
    CopyPool(const CopyPool&) = delete;
    CopyPool& operator=(const CopyPool&) = delete;
    void copy(void* dst, const void* src, size_t bytes);

private:
    void work();
    void loop();
    std::vector<std::thread> th_;
    size_t nworkers_ = 0;
    std::mutex job_mu_;  // one copy at a time
    std::mutex mu_;      // posting a copy; sleeping workers wait on cv_
    std::condition_variable cv_;
    char* dst_ = nullptr;
    const char* src_ = nullptr;
    size_t bytes_ = 0, chunk_ = 0, nchunks_ = 0;
    std::atomic<size_t> next_{0};
    std::atomic<size_t> finished_{0};  // workers done with the current copy
    std::atomic<uint64_t> gen_{0};     // copies posted
    std::atomic<bool> stop_{false};
};

// Host <-> device copies through two pinned chunks on a stream of its own:
// the host copies chunk i + 1 into (or out of) pinned memory while the DMA
// engine moves chunk i. Synchronous; `after` (optional) is an event the DMA
// waits for first (the producer of the data, or the last reader of the
// destination). One transfer at a time per stager (its own lock).
struct Stager {
    std::mutex mu;
    hipStream_t stream = nullptr;
    hipEvent_t ev[2] = {nullptr, nullptr};
    char* pin[2] = {nullptr, nullptr};
    size_t cap = 0;  // bytes of each pinned chunk
    CopyPool* pool = nullptr;
    bool ready = false;
    void init(CopyPool* p);
    void upload(void* d_dst, const void* h_src, size_t bytes, hipEvent_t after);
    void download(void* h_dst, const void* d_src, size_t bytes, hipEvent_t after);
    ~Stager();
};

// Pinned slab buffers of the pipelined host-buffer paths (hostpath.cpp):
// two for inputs, three for outputs (a slab's output is still being copied
// out while the next two compute).
struct HostSlabs {
    char* in[2] = {nullptr, nullptr};
    char* out[3] = {nullptr, nullptr, nullptr};
    char* sino[3] = {nullptr, nullptr, nullptr};
    size_t in_cap = 0, out_cap = 0, sino_cap = 0;
    ~HostSlabs();
};

// Most DMA pieces one slab download is split into (each with an event, so
// the host copies a piece out while the next one crosses PCIe).
constexpr int kSlabChunks = 8;

// Largest number of volumes in one batch.
constexpr int kMaxBatch = 256;
// Distinct row filter classes in one batch: the LDS FFT lengths rows take
// (64, 128, 256 for up to 90 bins; 2048, 4096 past 512 bins) and the three
// wave-form lengths (512 R, R = 1..3, for the rest).
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
    Buf rmax;  // per-row maxima of the FP8 GEMM filter form (grow-only)
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
    Buf vol, row, col, rflag, side, sino, filt, out, keys, mm, ell, pc_cnt, pc_pts, pc_coord, met;
    // cached device state
    std::map<std::pair<int, int>, float*> filters;  // (nfft, type) -> H[nfft/2+1]
    std::map<int, float2*> twiddles;                // nfft -> tw[nfft-1], stage order
    // wave-form row filter tables, (bins, type) -> one allocation: twiddles
    // (kWaveTabEntries float2), then the filter (kWaveHsFloats floats); bins
    // <= kWaveM and six types bound it to 3072 entries of 19 KB
    std::map<std::pair<int, int>, float2*> wave_tabs;
    // LDS FFT taps at a length shorter than Serial's nfft, (bins, type) -> Hs
    std::map<std::pair<int, int>, float*> short_taps;
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
    // Stage events: before the Radon stage, before the filter, before the
    // backprojection, after it; and for the Radon stage's breakdown, before
    // the min/max pass, before the normalisation, before the walk over the
    // image, between the two walks, after the walk over the transpose. Three
    // sets, so a pipelined slab records into one while the two before it are
    // still to be read; ev / evr point at the current set.
    hipEvent_t evs[3][4] = {};
    hipEvent_t evrs[3][5] = {};
    hipEvent_t* ev = evs[0];
    hipEvent_t* evr = evrs[0];
    double timings[4] = {0, 0, 0, 0};
    // Radon stage breakdown summed over the call's slabs: min/max, normalise,
    // walk over the image, walk over the transpose; and the slab count
    double radon_ms[4] = {0, 0, 0, 0};
    int radon_slabs = 0;
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
    // host-buffer paths: copy streams, memcpy threads, pinned staging
    hipStream_t h2d = nullptr, d2h = nullptr;
    // slab pipeline events (no timing): upload done and compute done per input
    // buffer, download done per output buffer and per piece of it
    hipEvent_t up_done[2] = {}, comp_done[2] = {}, dl_done[3] = {};
    hipEvent_t dl_piece[3][cthost::kSlabChunks] = {};
    hipEvent_t mark = nullptr;  // context stream position a stager transfer waits for
    std::unique_ptr<cthost::CopyPool> copier;
    cthost::Stager stager;
    cthost::HostSlabs slabs;
    Buf pin_vol[2], pout_vol[3], psino[3];  // device slab buffers of the pipeline
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
void check_radon(int w, int h, int bins);
void trim_trig(ct_hip_ctx* c);
void set_trig(ct_hip_ctx* c, const float* angles_deg, int na);
std::vector<float> serial_angles(int na);
size_t nzb_of(int nz);
void stage_radon(ct_hip_ctx* c, const float* d_vol, int w, int h, int nz, int na, int bins,
                 float* d_sino, float* d_mm);
void stage_fbp(ct_hip_ctx* c, const float* d_sino, int nz, int na, int bins, int N, int filter,
               const float* d_mm, float* d_out);
int bp_config_for(ct_hip_ctx* c, int N, int na);
float deg2rad(float d);
// Takes the allocation-free entry's scratch address on `device` (current).
void noalloc_init(int device);
// Adds the stage times of the events in the current set to c->timings: the
// Radon stage when with_radon, the filter and backprojection when with_fbp.
void add_timings(ct_hip_ctx* c, bool with_radon, bool with_fbp);
// The same for an explicit event set (ev: 4 stage events, evr: 5 Radon events).
void add_timings(ct_hip_ctx* c, const hipEvent_t* ev, const hipEvent_t* evr, bool with_radon,
                 bool with_fbp);
// Sizes the workspaces of the Radon stage (with_radon) and of the filter and
// backprojection stage (with_fbp) for nz slices, so later calls over at most
// nz slices reallocate nothing (a reallocation waits for the context stream).
void reserve_stages(ct_hip_ctx* c, int w, int h, int nz, int na, int bins, bool with_radon,
                    bool with_fbp);
void zero_timings(ct_hip_ctx* c);
// The context's memcpy pool and stager, created on first use.
CopyPool* copier(ct_hip_ctx* c);
Stager& stager(ct_hip_ctx* c);
// Pipelined host-buffer paths (hostpath.cpp).
void volume_host(ct_hip_ctx* c, const float* vol, int w, int h, int d, int na, int filter,
                 float* out, float* sino_out);
void from_sinos_host(ct_hip_ctx* c, const float* sinos, int nz, int na, int bins,
                     const float* angles_deg, int filter, float* out);
void release_host_paths(ct_hip_ctx* c);
// Device bytes of the volume path's workspace per slice: the normalised
// copies, the filtered sinogram, and the sinogram itself when sino_ws.
size_t volume_ws_per_slice(int w, int h, int na, bool sino_ws);
ct_hip_ctx::Trig get_trig(ct_hip_ctx* c, const float* angles_deg, int na);
float* get_filter(ct_hip_ctx* c, int nfft, int type);
float2* get_twiddles(ct_hip_ctx* c, int nfft);
// The row filter's arguments for one stack: the wave form (short transforms,
// rows of up to kWaveM bins) or the LDS FFT form, with their cached tables.
cthip::FiltArgs filter_args(ct_hip_ctx* c, const float* d_sino, int nz, int na, int bins, int filter,
                     float4* filt);
// Host copies of the wave form's tables (kWaveTabEntries float2, kWaveHsFloats
// floats); R = wave_filter_r(...) > 0 for this geometry.
void wave_tables_host(const Geom& g, int filter, float2* tw, float* hs);
// Transform length of the LDS FFT form: the smallest power of two (>= 64)
// that holds the kernel's offsets, 2D + 1, or Serial's nfft if that is not
// longer; lds_taps_host gives its L/2 + 1 taps (Serial's own at nfft).
int lds_filter_length(const Geom& g);
void lds_taps_host(const Geom& g, int filter, int L, float* H);
hipError_t launch_filter(const cthip::FiltArgs& p, hipStream_t s);
// Toeplitz tables of the GEMM row filter forms, cached per (nfft, type, bins)
// and made on the context stream (which the call drains).
const ct_hip_ctx::Toep& get_toeplitz(ct_hip_ctx* c, const Geom& g, int type);
// The context backend's Toeplitz-GEMM row filter of one stack on `stream`;
// the FP8 form needs nz * na floats of row maxima in rmax.
void launch_gemm_filter(ct_hip_ctx* c, const float* d_sino, int nz, int na, int bins, int filter,
                        float4* filt, float* rmax, hipStream_t stream);
// Work-list class of a row filter launch: nfft for the LDS form, -R for the
// wave form (a persistent launch holds one class).
inline int filter_class(const cthip::FiltArgs& p) { return p.wave_r > 0 ? -p.wave_r : p.nfft; }
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
