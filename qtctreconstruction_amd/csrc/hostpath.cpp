// hostpath.cpp — host-buffer transfers and the pipelined host-buffer entry
// points: reconstructVolume and reconstructVolumeFromSinograms on caller host
// memory (IReconstructionBackend.h:28-41), the counterpart of the reference
// CUDA path's pinned staging and ping-pong streams (CUDABackend.cu:405-420,
// 736-863).
//
// On this part pageable memory crosses PCIe at ~18 GB/s, pinned memory at
// ~57 GB/s each way (~97 GB/s both ways at once), and a few host threads copy
// pageable <-> pinned at ~100 GB/s. So:
//   * every host transfer goes through pinned staging, the host copy of one
//     chunk overlapping the DMA of the next (Stager);
//   * the volume entry points run the volume in z-slabs on three streams: the
//     upload of slab k+1 and the download of slab k-1 overlap the compute of
//     slab k. The calling thread copies inputs into pinned memory and
//     enqueues uploads, compute and downloads; a second thread copies the
//     outputs out, piece by piece as the DMA delivers them, so no upload
//     waits behind a copy-out. Both use a small memcpy pool (CopyPool).
//     Device memory holds two input and three output slabs and one slab's
//     workspace, not the whole volume's;
//   * the slabs ramp up and down (the first and last small, so little
//     transfer time is left outside the compute; the middle ones large, as
//     every launch over few slices wastes part of the chip).
// Slab boundaries are multiples of 8 slices, so every Radon z-block and every
// pair of rows the FFT filter transforms together is the one of the
// whole-volume device call, and the results are the same bit for bit.
#include "ctx.h"

#include <algorithm>
#include <chrono>
#include <cstring>
#include <exception>

using namespace cthip;

namespace cthost {

// ------------------------------------------------------------ CopyPool
constexpr size_t kCopyMinParallel = size_t(1) << 20;  // below: the caller alone

CopyPool::CopyPool(int workers) {
    nworkers_ = (size_t)std::max(workers, 0);
    for (size_t i = 0; i < nworkers_; ++i) th_.emplace_back([this] { loop(); });
}

CopyPool::~CopyPool() {
    {
        std::lock_guard<std::mutex> lk(mu_);
        stop_.store(true);
    }
    cv_.notify_all();
    for (auto& t : th_) t.join();
}

void CopyPool::work() {
    for (;;) {
        const size_t i = next_.fetch_add(1, std::memory_order_relaxed);
        if (i >= nchunks_) break;
        const size_t off = i * chunk_;
        std::memcpy(dst_ + off, src_ + off, std::min(chunk_, bytes_ - off));
    }
}

// Spin-wait step: a pause where the host has one.
static inline void spin_pause() {
#if defined(__x86_64__)
    __builtin_ia32_pause();
#else
    std::this_thread::yield();
#endif
}

// How long an idle worker keeps polling for the next copy before it sleeps:
// a pipelined call asks for copies a fraction of a millisecond apart, and a
// wake-up from sleep costs tens of microseconds per copy.
constexpr auto kCopySpin = std::chrono::milliseconds(2);

// A worker joins every copy (the caller waits for all of them to finish
// theirs), so none is left holding the fields of a finished copy, and none
// can miss a copy: the next is posted only after every worker has finished
// this one.
void CopyPool::loop() {
    uint64_t seen = 0;
    for (;;) {
        const auto t0 = std::chrono::steady_clock::now();
        while (gen_.load(std::memory_order_acquire) == seen &&
               !stop_.load(std::memory_order_relaxed) &&
               std::chrono::steady_clock::now() - t0 < kCopySpin)
            spin_pause();
        {
            std::unique_lock<std::mutex> lk(mu_);
            cv_.wait(lk, [&] { return stop_.load() || gen_.load() != seen; });
            if (stop_.load()) return;
            seen = gen_.load();
        }
        work();
        finished_.fetch_add(1, std::memory_order_release);
    }
}

void CopyPool::copy(void* dst, const void* src, size_t bytes) {
    if (bytes == 0) return;
    if (bytes < kCopyMinParallel || nworkers_ == 0) {
        std::memcpy(dst, src, bytes);
        return;
    }
    std::lock_guard<std::mutex> job(job_mu_);
    {
        std::lock_guard<std::mutex> lk(mu_);
        dst_ = static_cast<char*>(dst);
        src_ = static_cast<const char*>(src);
        bytes_ = bytes;
        // about two tasks per thread, 256 KB to 2 MB each
        const size_t per = bytes / (2 * (nworkers_ + 1));
        chunk_ = std::min(std::max(per, size_t(256) << 10), size_t(2) << 20) & ~size_t(4095);
        nchunks_ = (bytes + chunk_ - 1) / chunk_;
        next_.store(0, std::memory_order_relaxed);
        finished_.store(0, std::memory_order_relaxed);
        gen_.fetch_add(1, std::memory_order_release);
    }
    cv_.notify_all();
    work();
    // the workers are awake or about to be: poll rather than sleep
    for (unsigned n = 0; finished_.load(std::memory_order_acquire) < nworkers_; ++n)
        if (n < 4096)
            spin_pause();
        else
            std::this_thread::yield();
}

// The caller plus up to 7 workers: 8 threads reached ~110 GB/s on the host
// of this part, 16 no more.
CopyPool* copier(ct_hip_ctx* c) {
    if (!c->copier) {
        const unsigned hw = std::max(2u, std::thread::hardware_concurrency());
        const int total = (int)std::min(8u, std::max(2u, hw / 2));
        c->copier.reset(new CopyPool(total - 1));
    }
    return c->copier.get();
}

// ------------------------------------------------------------ Stager
constexpr size_t kStageChunk = size_t(32) << 20;

static void pinned_alloc(char** p, size_t bytes) {
    if (hipHostMalloc(reinterpret_cast<void**>(p), bytes, hipHostMallocDefault) != hipSuccess ||
        !*p) {
        (void)hipGetLastError();
        *p = nullptr;
        throw StatusError(CT_ERR_NOMEM, "pinned host allocation of " + std::to_string(bytes) +
                                            " B failed");
    }
}

void Stager::init(CopyPool* p) {
    if (ready) return;
    pool = p;
    hip_check(hipStreamCreateWithFlags(&stream, hipStreamNonBlocking), "stager stream");
    for (auto& e : ev) hip_check(hipEventCreateWithFlags(&e, hipEventDisableTiming), "stager event");
    for (auto& b : pin) pinned_alloc(&b, kStageChunk);
    cap = kStageChunk;
    ready = true;
}

Stager::~Stager() {
    if (stream) (void)hipStreamSynchronize(stream);
    for (auto& b : pin)
        if (b) (void)hipHostFree(b);
    for (auto& e : ev)
        if (e) (void)hipEventDestroy(e);
    if (stream) (void)hipStreamDestroy(stream);
}

void Stager::upload(void* d_dst, const void* h_src, size_t bytes, hipEvent_t after) {
    std::lock_guard<std::mutex> lk(mu);
    if (after) hip_check(hipStreamWaitEvent(stream, after, 0), "stager wait");
    char* d = static_cast<char*>(d_dst);
    const char* h = static_cast<const char*>(h_src);
    size_t i = 0;
    for (size_t off = 0; off < bytes; off += cap, ++i) {
        const size_t n = std::min(cap, bytes - off);
        const int b = (int)(i & 1);
        // pin[b] is free once the DMA of chunk i - 2 has read it
        if (i >= 2) hip_check(hipEventSynchronize(ev[b]), "stager sync");
        pool->copy(pin[b], h + off, n);
        hip_check(hipMemcpyAsync(d + off, pin[b], n, hipMemcpyHostToDevice, stream), "upload");
        hip_check(hipEventRecord(ev[b], stream), "stager event");
    }
    hip_check(hipStreamSynchronize(stream), "upload sync");
}

void Stager::download(void* h_dst, const void* d_src, size_t bytes, hipEvent_t after) {
    std::lock_guard<std::mutex> lk(mu);
    if (bytes == 0) return;
    if (after) hip_check(hipStreamWaitEvent(stream, after, 0), "stager wait");
    char* h = static_cast<char*>(h_dst);
    const char* d = static_cast<const char*>(d_src);
    const size_t nc = (bytes + cap - 1) / cap;
    auto enqueue = [&](size_t i) {
        const size_t off = i * cap, n = std::min(cap, bytes - off);
        const int b = (int)(i & 1);
        hip_check(hipMemcpyAsync(pin[b], d + off, n, hipMemcpyDeviceToHost, stream), "download");
        hip_check(hipEventRecord(ev[b], stream), "stager event");
    };
    enqueue(0);
    for (size_t i = 0; i < nc; ++i) {
        // pin[(i + 1) & 1] held chunk i - 1, copied out in the last round
        if (i + 1 < nc) enqueue(i + 1);
        hip_check(hipEventSynchronize(ev[i & 1]), "stager sync");
        const size_t off = i * cap;
        pool->copy(h + off, pin[i & 1], std::min(cap, bytes - off));
    }
}

Stager& stager(ct_hip_ctx* c) {
    c->stager.init(copier(c));
    return c->stager;
}


// ------------------------------------------------------------ slabs
template <size_t N>
static void free_pinned(char* (&b)[N]) {
    for (auto& x : b) {
        if (x) (void)hipHostFree(x);
        x = nullptr;
    }
}

template <size_t N>
static void grow_pinned(char* (&b)[N], size_t& cap, size_t bytes) {
    if (cap >= bytes) return;
    free_pinned(b);
    cap = 0;
    for (auto& x : b) pinned_alloc(&x, bytes);
    cap = bytes;
}

HostSlabs::~HostSlabs() {
    free_pinned(in);
    free_pinned(out);
    free_pinned(sino);
}

void release_host_paths(ct_hip_ctx* c) {
    for (Buf& b : c->pin_vol) release(c, b);
    for (Buf& b : c->pout_vol) release(c, b);
    for (Buf& b : c->psino) release(c, b);
    HostSlabs& s = c->slabs;
    free_pinned(s.in);
    free_pinned(s.out);
    free_pinned(s.sino);
    s.in_cap = s.out_cap = s.sino_cap = 0;
}

size_t volume_ws_per_slice(int w, int h, int na, bool sino_ws) {
    const size_t sq = (size_t)geometry(w).sq;
    return (radon_zbytes(w, h) + radon_zbytes(h, w)) / kRZ +
           2 * sizeof(int) * ((size_t)h + w) / kRZ +  // row and column extents
           (size_t)na * sq * sizeof(float) +
           (sino_ws ? (size_t)na * w * sizeof(float) : 0);
}

// Slab sizes of the pipelined paths: `first` for the two ends, whose upload
// (first slab) and download (last) cannot hide behind compute, about 32 MB of
// input; `most` for the middle, four times that, since a launch over few
// slices leaves part of the chip idle (the plain path at 512 x 512 slices and
// 720 angles: 58 us per slice over 32 slices, 47.5 over 64, 46.2 over 128,
// 43.5 over 512). Multiples of 8 slices. With a budget, the device slabs
// and the workspace of the largest slab must fit it.
struct SlabSizes {
    int first, most;
};

static SlabSizes host_slab(ct_hip_ctx* c, int d, size_t slice_bytes, size_t dev_per_slice,
                           std::initializer_list<const Buf*> reused) {
    size_t F = (size_t(32) << 20) / std::max<size_t>(slice_bytes, 1);
    F = std::max<size_t>((F + kRZ - 1) / kRZ * kRZ, kRZ);
    size_t M = 4 * F;
    if (c->budget) {
        size_t held = 0;  // workspaces the call reuses count as available
        for (const Buf* b : reused) held += b->cap;
        const size_t used = c->in_use - held;
        const size_t avail = c->budget > used ? c->budget - used : 0;
        const size_t fit = avail / std::max<size_t>(dev_per_slice, 1) * 15 / 16;
        const size_t least = (size_t)std::min(d, kRZ);
        if (fit < least)
            throw StatusError(CT_ERR_BUDGET, "the HBM budget holds no slab of " +
                                                 std::to_string(least) + " slices");
        if (fit < M) M = fit >= (size_t)kRZ ? fit / kRZ * kRZ : fit;  // fit < 8 only when d < 8
        F = std::min(F, M);
    }
    return {(int)F, (int)M};
}

// The slab sequence over d slices: a ramp F, 2F, 4F, ... up to M at the
// start, slabs of the ramp's top size in the middle, the ramp reversed at the
// end, so that each slab's transfers fit inside its neighbours' compute. Every
// slab but the last starts and ends on a multiple of 8 slices, and none
// exceeds M.
static std::vector<int> slab_plan(int d, int F, int M) {
    const int d8 = d / kRZ * kRZ, ragged = d - d8;
    if (d8 == 0) return {d};
    std::vector<int> R{F};
    while (R.back() < M) R.push_back(std::min(2 * R.back(), M));
    // the longest ramp whose two ends and one slab of the next size fit
    size_t k = 0;
    long ramp = 0;  // slices in R[0 .. k-1]
    while (k + 1 < R.size() && 2 * (ramp + R[k]) + R[k + 1] <= d8) ramp += R[k++];
    const int C = R[k];
    const long mid = d8 - 2 * ramp;
    const int r = (int)(mid % C);
    std::vector<int> plan(R.begin(), R.begin() + (long)k);
    for (long q = mid / C; q > 0; --q) plan.push_back(C);
    if (r && 2 * r >= C) plan.push_back(r);
    for (size_t i = k; i-- > 0;) plan.push_back(R[i]);
    if (r && 2 * r < C) plan.push_back(r);  // a small rest goes last
    if (ragged) {
        if (plan.back() + ragged <= M)
            plan.back() += ragged;
        else
            plan.push_back(ragged);
    }
    return plan;
}

// Buffers and byte sizes of one pipelined call: per slab, the host input
// (in_sl bytes per slice) goes up into a device input slab, the compute turns
// it into a device output slab (out_sl bytes per slice) and, when sdst is
// set, a sinogram slab (sino_sl), and both come back into the host outputs.
struct SlabIO {
    const char* src;
    size_t in_sl;
    char* dst;
    size_t out_sl;
    char* sdst;
    size_t sino_sl;
    char* din[2];
    char* dout[3];
    char* dsino[3];
    bool radon, fbp;  // the stage times each slab records
};

// Orders stream s after event e, unless e has completed already (a wait puts
// a barrier packet on the stream even then).
static void stream_after(hipStream_t s, hipEvent_t e) {
    const hipError_t q = hipEventQuery(e);
    if (q == hipSuccess) return;
    if (q != hipErrorNotReady) hip_check(q, "event query");
    hip_check(hipStreamWaitEvent(s, e, 0), "slab wait");
}

// Bytes per DMA piece of a slab transfer: an eighth of it, at least 4 MB.
static size_t piece_bytes(size_t bytes) {
    const size_t p = (bytes + kSlabChunks - 1) / kSlabChunks;
    return std::max<size_t>((p + 4095) & ~size_t(4095), size_t(4) << 20);
}

// Runs the slabs of `plan`. The calling thread copies slab j into pinned
// input j % 2 and enqueues its upload (h2d), compute(j, nz, in, out, sino)
// (context stream) and its download into pinned output j % 3 (d2h), piece
// by piece with an event each; an output thread copies each slab out as its
// pieces land and adds its stage times. Waits: an upload for the compute
// that last read its device input (slab j - 2), a compute for its upload and
// for the download that last read its device output (slab j - 3), a
// download for its compute. The calling thread waits for the upload of slab
// j - 2 before reusing its pinned input, and for slab j - 3 to be copied out
// before reusing its pinned output and stage events. On an error the output
// thread is stopped and every stream drained before the buffers can be
// reused.
template <typename Compute>
static void run_slabs(ct_hip_ctx* c, const std::vector<int>& plan, const SlabIO& io,
                      Compute&& compute) {
    const int K = (int)plan.size();
    std::vector<size_t> z0(K + 1, 0);
    for (int k = 0; k < K; ++k) z0[k + 1] = z0[k] + (size_t)plan[k];
    CopyPool* cp = copier(c);
    HostSlabs& hs = c->slabs;
    std::mutex mu;
    std::condition_variable cv;
    int issued = 0, done = 0;
    bool stop = false;
    std::exception_ptr out_err;
    std::thread out([&] {
        try {
            hip_check(hipSetDevice(c->device), "hipSetDevice");
            for (int k = 0; k < K; ++k) {
                {
                    std::unique_lock<std::mutex> lk(mu);
                    cv.wait(lk, [&] { return stop || issued > k; });
                    if (stop) return;
                }
                const int t = k % 3;
                const size_t ob = plan[k] * io.out_sl, pb = piece_bytes(ob);
                for (size_t off = 0, i = 0; off < ob; off += pb, ++i) {
                    hip_check(hipEventSynchronize(c->dl_piece[t][i]), "slab sync");
                    cp->copy(io.dst + z0[k] * io.out_sl + off, hs.out[t] + off,
                             std::min(pb, ob - off));
                }
                hip_check(hipEventSynchronize(c->dl_done[t]), "slab sync");
                if (io.sdst)
                    cp->copy(io.sdst + z0[k] * io.sino_sl, hs.sino[t], plan[k] * io.sino_sl);
                add_timings(c, c->evs[t], c->evrs[t], io.radon, io.fbp);
                {
                    std::lock_guard<std::mutex> lk(mu);
                    done = k + 1;
                }
                cv.notify_all();
            }
        } catch (...) {
            std::lock_guard<std::mutex> lk(mu);
            out_err = std::current_exception();
            cv.notify_all();
        }
    });
    auto wait_done = [&](int n) {
        std::unique_lock<std::mutex> lk(mu);
        cv.wait(lk, [&] { return out_err || done >= n; });
        if (out_err) std::rethrow_exception(out_err);
    };
    try {
        for (int j = 0; j < K; ++j) {
            const int p = j & 1, t = j % 3;
            const size_t ib = plan[j] * io.in_sl, ipb = piece_bytes(ib);
            if (j >= 2) {
                hip_check(hipEventSynchronize(c->up_done[p]), "slab sync");
                stream_after(c->h2d, c->comp_done[p]);
            }
            for (size_t off = 0; off < ib; off += ipb) {
                const size_t n = std::min(ipb, ib - off);
                cp->copy(hs.in[p] + off, io.src + z0[j] * io.in_sl + off, n);
                hip_check(hipMemcpyAsync(io.din[p] + off, hs.in[p] + off, n,
                                         hipMemcpyHostToDevice, c->h2d),
                          "upload slab");
            }
            hip_check(hipEventRecord(c->up_done[p], c->h2d), "event");
            if (j >= 3) wait_done(j - 2);
            stream_after(c->stream, c->up_done[p]);
            if (j >= 3) stream_after(c->stream, c->dl_done[t]);
            c->ev = c->evs[t];
            c->evr = c->evrs[t];
            compute(plan[j], io.din[p], io.dout[t], io.dsino[t]);
            hip_check(hipEventRecord(c->comp_done[p], c->stream), "event");
            hip_check(hipStreamWaitEvent(c->d2h, c->comp_done[p], 0), "slab wait");
            const size_t ob = plan[j] * io.out_sl, opb = piece_bytes(ob);
            for (size_t off = 0, i = 0; off < ob; off += opb, ++i) {
                hip_check(hipMemcpyAsync(hs.out[t] + off, io.dout[t] + off,
                                         std::min(opb, ob - off), hipMemcpyDeviceToHost, c->d2h),
                          "download slab");
                hip_check(hipEventRecord(c->dl_piece[t][i], c->d2h), "event");
            }
            if (io.sdst)
                hip_check(hipMemcpyAsync(hs.sino[t], io.dsino[t], plan[j] * io.sino_sl,
                                         hipMemcpyDeviceToHost, c->d2h),
                          "download sinograms");
            hip_check(hipEventRecord(c->dl_done[t], c->d2h), "event");
            {
                std::lock_guard<std::mutex> lk(mu);
                issued = j + 1;
            }
            cv.notify_all();
        }
        wait_done(K);
    } catch (...) {
        {
            std::lock_guard<std::mutex> lk(mu);
            stop = true;
        }
        cv.notify_all();
        out.join();
        (void)hipStreamSynchronize(c->h2d);
        (void)hipStreamSynchronize(c->stream);
        (void)hipStreamSynchronize(c->d2h);
        c->ev = c->evs[0];
        c->evr = c->evrs[0];
        throw;
    }
    out.join();
    c->ev = c->evs[0];
    c->evr = c->evrs[0];
}

static int max_slab(const std::vector<int>& plan) {
    return *std::max_element(plan.begin(), plan.end());
}

// reconstructVolume on host buffers (SerialBackend.cpp:261-318): per slab,
// normalise + Radon (bins = w), filter + backprojection + denormalisation.
void volume_host(ct_hip_ctx* c, const float* vol, int w, int h, int d, int na, int filter,
                 float* out, float* sino_out) {
    const int bins = w;  // SerialBackend.cpp:287-288
    const bool fbp = (h == w);
    const std::vector<float> ang = serial_angles(na);
    set_trig(c, ang.data(), na);
    if (fbp) (void)bp_config_for(c, w, na);
    zero_timings(c);
    const size_t sl_in = (size_t)w * h * sizeof(float);
    const size_t sl_sino = (size_t)na * bins * sizeof(float);
    const size_t dev_per = 5 * sl_in + (sino_out ? 3 : 1) * sl_sino +
                           volume_ws_per_slice(w, h, na, false);
    const SlabSizes sz =
        host_slab(c, d, sl_in, dev_per,
                  {&c->pin_vol[0], &c->pin_vol[1], &c->pout_vol[0], &c->pout_vol[1],
                   &c->pout_vol[2], &c->psino[0], &c->psino[1], &c->psino[2], &c->row, &c->col,
                   &c->filt});
    const std::vector<int> plan = slab_plan(d, sz.first, sz.most);
    const size_t S = (size_t)max_slab(plan);
    grow_pinned(c->slabs.in, c->slabs.in_cap, S * sl_in);
    grow_pinned(c->slabs.out, c->slabs.out_cap, S * sl_in);
    if (sino_out) grow_pinned(c->slabs.sino, c->slabs.sino_cap, S * sl_sino);
    SlabIO io{reinterpret_cast<const char*>(vol), sl_in, reinterpret_cast<char*>(out), sl_in,
              reinterpret_cast<char*>(sino_out), sl_sino, {}, {}, {}, true, fbp};
    for (int p = 0; p < 2; ++p) io.din[p] = ensure<char>(c, c->pin_vol[p], S * sl_in);
    for (int t = 0; t < 3; ++t) {
        io.dout[t] = ensure<char>(c, c->pout_vol[t], S * sl_in);
        io.dsino[t] =
            (t == 0 || sino_out) ? ensure<char>(c, c->psino[t], S * sl_sino) : io.dsino[0];
    }
    float* d_mm = ensure<float>(c, c->mm, sizeof(float) * 2 * S);
    reserve_stages(c, w, h, (int)S, na, bins, true, fbp);
    run_slabs(c, plan, io, [&](int nz, char* in, char* o, char* sino) {
        float* fsino = reinterpret_cast<float*>(sino);
        hip_check(hipEventRecord(c->ev[0], c->stream), "event");
        stage_radon(c, reinterpret_cast<const float*>(in), w, h, nz, na, bins, fsino, d_mm);
        if (fbp) {
            stage_fbp(c, fsino, nz, na, bins, w, filter, d_mm, reinterpret_cast<float*>(o));
        } else {  // Buffer3D::setSlice drops slices of another shape (CTTypes.h:112-116)
            hip_check(hipEventRecord(c->ev[1], c->stream), "event");
            hip_check(hipMemsetAsync(o, 0, nz * sl_in, c->stream), "zero");
        }
    });
}

// reconstructVolumeFromSinograms on host buffers (BackendFactory.cpp:17-60):
// per slab, filter + backprojection onto bins x bins, no denormalisation.
void from_sinos_host(ct_hip_ctx* c, const float* sinos, int nz, int na, int bins,
                     const float* angles_deg, int filter, float* out) {
    set_trig(c, angles_deg, na);
    (void)bp_config_for(c, bins, na);
    zero_timings(c);
    const size_t sl_in = (size_t)na * bins * sizeof(float);
    const size_t sl_out = (size_t)bins * bins * sizeof(float);
    const size_t dev_per = 2 * sl_in + 3 * sl_out + (size_t)na * geometry(bins).sq * sizeof(float);
    const SlabSizes sz = host_slab(c, nz, std::max(sl_in, sl_out), dev_per,
                                   {&c->pin_vol[0], &c->pin_vol[1], &c->pout_vol[0],
                                    &c->pout_vol[1], &c->pout_vol[2], &c->filt});
    const std::vector<int> plan = slab_plan(nz, sz.first, sz.most);
    const size_t S = (size_t)max_slab(plan);
    grow_pinned(c->slabs.in, c->slabs.in_cap, S * sl_in);
    grow_pinned(c->slabs.out, c->slabs.out_cap, S * sl_out);
    SlabIO io{reinterpret_cast<const char*>(sinos), sl_in, reinterpret_cast<char*>(out), sl_out,
              nullptr, 0, {}, {}, {}, false, true};
    for (int p = 0; p < 2; ++p) io.din[p] = ensure<char>(c, c->pin_vol[p], S * sl_in);
    for (int t = 0; t < 3; ++t) io.dout[t] = ensure<char>(c, c->pout_vol[t], S * sl_out);
    reserve_stages(c, bins, bins, (int)S, na, bins, false, true);
    run_slabs(c, plan, io, [&](int n, char* in, char* o, char*) {
        stage_fbp(c, reinterpret_cast<const float*>(in), n, na, bins, bins, filter, nullptr,
                  reinterpret_cast<float*>(o));
    });
}

}  // namespace cthost
