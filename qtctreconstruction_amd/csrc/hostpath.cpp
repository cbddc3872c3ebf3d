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
//     slab k, while the calling thread and a small thread pool (CopyPool)
//     copy slab k+1 into pinned memory and slab k-1 out of it. Device memory
//     then holds two input and two output slabs and one slab's workspace,
//     not the whole volume's.
// Slab boundaries are multiples of 8 slices, so every Radon z-block and every
// pair of rows the FFT filter transforms together is the one of the
// whole-volume device call, and the results are the same bit for bit.
#include "ctx.h"

#include <algorithm>
#include <cstring>

using namespace cthip;

namespace cthost {

// ------------------------------------------------------------ CopyPool
constexpr size_t kCopyChunk = size_t(2) << 20;        // bytes per task
constexpr size_t kCopyMinParallel = size_t(4) << 20;  // below: the caller alone

CopyPool::CopyPool(int workers) {
    nworkers_ = (size_t)std::max(workers, 0);
    for (size_t i = 0; i < nworkers_; ++i) th_.emplace_back([this] { loop(); });
}

CopyPool::~CopyPool() {
    {
        std::lock_guard<std::mutex> lk(mu_);
        stop_ = true;
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

// A worker joins every copy (the caller waits for all of them to finish
// theirs), so none is left holding the fields of a finished copy.
void CopyPool::loop() {
    uint64_t seen = 0;
    for (;;) {
        {
            std::unique_lock<std::mutex> lk(mu_);
            cv_.wait(lk, [&] { return stop_ || gen_ != seen; });
            if (stop_) return;
            seen = gen_;
        }
        work();
        std::lock_guard<std::mutex> lk(mu_);
        if (++finished_ == nworkers_) cv_done_.notify_one();
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
        chunk_ = kCopyChunk;
        nchunks_ = (bytes + kCopyChunk - 1) / kCopyChunk;
        next_.store(0, std::memory_order_relaxed);
        finished_ = 0;
        ++gen_;
    }
    cv_.notify_all();
    work();
    std::unique_lock<std::mutex> lk(mu_);
    cv_done_.wait(lk, [&] { return finished_ == nworkers_; });
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
HostSlabs::~HostSlabs() {
    for (char** pair : {in, out, sino})
        for (int k = 0; k < 2; ++k)
            if (pair[k]) (void)hipHostFree(pair[k]);
}

static void grow_pinned(char* (&b)[2], size_t& cap, size_t bytes) {
    if (cap >= bytes) return;
    for (auto& x : b) {
        if (x) (void)hipHostFree(x);
        x = nullptr;
    }
    cap = 0;
    for (auto& x : b) pinned_alloc(&x, bytes);
    cap = bytes;
}

void release_host_paths(ct_hip_ctx* c) {
    for (int k = 0; k < 2; ++k) {
        release(c, c->pin_vol[k]);
        release(c, c->pout_vol[k]);
        release(c, c->psino[k]);
    }
    HostSlabs& s = c->slabs;
    for (char** pair : {s.in, s.out, s.sino})
        for (int k = 0; k < 2; ++k) {
            if (pair[k]) (void)hipHostFree(pair[k]);
            pair[k] = nullptr;
        }
    s.in_cap = s.out_cap = s.sino_cap = 0;
}

size_t volume_ws_per_slice(int w, int h, int na, bool sino_ws) {
    const size_t sq = (size_t)geometry(w).sq;
    return (radon_zbytes(w, h) + radon_zbytes(h, w)) / kRZ + (size_t)na * sq * sizeof(float) +
           (sino_ws ? (size_t)na * w * sizeof(float) : 0);
}

// Slices per slab of the pipelined paths: about an eighth of the volume, so
// transfers hide behind compute except for the first upload and the last
// download, and at least 16 MB of input, so each slab's launches stay large;
// a multiple of 8. With a budget, the device slabs and workspace must fit it.
static int host_slab(ct_hip_ctx* c, int d, size_t in_slice, size_t dev_per_slice,
                     std::initializer_list<const Buf*> reused) {
    size_t S = std::max<size_t>((d + 7) / 8, (size_t(16) << 20) / std::max<size_t>(in_slice, 1));
    S = (S + kRZ - 1) / kRZ * kRZ;
    if (S >= (size_t)d) S = (size_t)d;
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
        if (fit < S) S = fit / kRZ * kRZ;
    }
    return (int)S;
}

// Runs slabs 0 .. K-1: launch(k) enqueues slab k's upload, compute and
// download; finish(k) waits for slab k's download and copies it out. Slab k
// is launched before slab k-1 is finished, so the device never idles on the
// host copies. On an error every stream is drained before the buffers can be
// reused.
template <typename L, typename F>
static void run_slabs(ct_hip_ctx* c, int K, L&& launch, F&& finish) {
    try {
        for (int k = 0; k < K; ++k) {
            launch(k);
            if (k >= 1) finish(k - 1);
        }
        finish(K - 1);
    } catch (...) {
        (void)hipStreamSynchronize(c->h2d);
        (void)hipStreamSynchronize(c->stream);
        (void)hipStreamSynchronize(c->d2h);
        c->ev = c->evs[0];
        c->evr = c->evrs[0];
        throw;
    }
    c->ev = c->evs[0];
    c->evr = c->evrs[0];
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
    const size_t wh = (size_t)w * h;
    const size_t sl_in = wh * sizeof(float), sl_sino = (size_t)na * bins * sizeof(float);
    const size_t dev_per = 4 * sl_in + (sino_out ? 2 : 1) * sl_sino +
                           volume_ws_per_slice(w, h, na, false);
    const int S = host_slab(c, d, sl_in, dev_per,
                            {&c->pin_vol[0], &c->pin_vol[1], &c->pout_vol[0], &c->pout_vol[1],
                             &c->psino[0], &c->psino[1], &c->row, &c->col, &c->filt});
    const int K = (d + S - 1) / S;
    CopyPool* cp = copier(c);
    grow_pinned(c->slabs.in, c->slabs.in_cap, S * sl_in);
    grow_pinned(c->slabs.out, c->slabs.out_cap, S * sl_in);
    if (sino_out) grow_pinned(c->slabs.sino, c->slabs.sino_cap, S * sl_sino);
    float *din[2], *dout[2], *dsino[2];
    for (int p = 0; p < 2; ++p) {
        din[p] = ensure<float>(c, c->pin_vol[p], S * sl_in);
        dout[p] = ensure<float>(c, c->pout_vol[p], S * sl_in);
    }
    dsino[0] = ensure<float>(c, c->psino[0], S * sl_sino);
    dsino[1] = sino_out ? ensure<float>(c, c->psino[1], S * sl_sino) : dsino[0];
    float* d_mm = ensure<float>(c, c->mm, sizeof(float) * 2 * (size_t)S);
    auto launch = [&](int k) {
        const int p = k & 1, z0 = k * S, nz = std::min(S, d - z0);
        // pinned input p is free once the upload of slab k-2 has read it
        if (k >= 2) hip_check(hipEventSynchronize(c->pev[0][p]), "slab sync");
        cp->copy(c->slabs.in[p], vol + z0 * wh, nz * sl_in);
        // device input p is free once slab k-2's compute has read it
        if (k >= 2) hip_check(hipStreamWaitEvent(c->h2d, c->pev[1][p], 0), "slab wait");
        hip_check(hipMemcpyAsync(din[p], c->slabs.in[p], nz * sl_in, hipMemcpyHostToDevice, c->h2d),
                  "upload slab");
        hip_check(hipEventRecord(c->pev[0][p], c->h2d), "event");
        // compute after the upload; device output p is free once slab k-2's
        // download has read it
        hip_check(hipStreamWaitEvent(c->stream, c->pev[0][p], 0), "slab wait");
        if (k >= 2) hip_check(hipStreamWaitEvent(c->stream, c->pev[2][p], 0), "slab wait");
        c->ev = c->evs[p];
        c->evr = c->evrs[p];
        hip_check(hipEventRecord(c->ev[0], c->stream), "event");
        stage_radon(c, din[p], w, h, nz, na, bins, dsino[p], d_mm);
        if (fbp) {
            stage_fbp(c, dsino[p], nz, na, bins, w, filter, d_mm, dout[p]);
        } else {  // Buffer3D::setSlice drops slices of another shape (CTTypes.h:112-116)
            hip_check(hipEventRecord(c->ev[1], c->stream), "event");
            hip_check(hipMemsetAsync(dout[p], 0, nz * sl_in, c->stream), "zero");
        }
        hip_check(hipEventRecord(c->pev[1][p], c->stream), "event");
        hip_check(hipStreamWaitEvent(c->d2h, c->pev[1][p], 0), "slab wait");
        hip_check(hipMemcpyAsync(c->slabs.out[p], dout[p], nz * sl_in, hipMemcpyDeviceToHost,
                                 c->d2h),
                  "download slab");
        if (sino_out)
            hip_check(hipMemcpyAsync(c->slabs.sino[p], dsino[p], nz * sl_sino,
                                     hipMemcpyDeviceToHost, c->d2h),
                      "download sinograms");
        hip_check(hipEventRecord(c->pev[2][p], c->d2h), "event");
    };
    auto finish = [&](int k) {
        const int p = k & 1, z0 = k * S, nz = std::min(S, d - z0);
        hip_check(hipEventSynchronize(c->pev[2][p]), "slab sync");
        c->ev = c->evs[p];
        c->evr = c->evrs[p];
        add_timings(c, true, fbp);
        cp->copy(out + z0 * wh, c->slabs.out[p], nz * sl_in);
        if (sino_out)
            cp->copy(sino_out + (size_t)z0 * na * bins, c->slabs.sino[p], nz * sl_sino);
    };
    run_slabs(c, K, launch, finish);
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
    const size_t dev_per = 2 * sl_in + 2 * sl_out + (size_t)na * geometry(bins).sq * sizeof(float);
    const int S = host_slab(c, nz, std::max(sl_in, sl_out), dev_per,
                            {&c->pin_vol[0], &c->pin_vol[1], &c->pout_vol[0], &c->pout_vol[1],
                             &c->filt});
    const int K = (nz + S - 1) / S;
    CopyPool* cp = copier(c);
    grow_pinned(c->slabs.in, c->slabs.in_cap, S * sl_in);
    grow_pinned(c->slabs.out, c->slabs.out_cap, S * sl_out);
    float *din[2], *dout[2];
    for (int p = 0; p < 2; ++p) {
        din[p] = ensure<float>(c, c->pin_vol[p], S * sl_in);
        dout[p] = ensure<float>(c, c->pout_vol[p], S * sl_out);
    }
    auto launch = [&](int k) {
        const int p = k & 1, z0 = k * S, n = std::min(S, nz - z0);
        if (k >= 2) hip_check(hipEventSynchronize(c->pev[0][p]), "slab sync");
        cp->copy(c->slabs.in[p], sinos + (size_t)z0 * na * bins, n * sl_in);
        if (k >= 2) hip_check(hipStreamWaitEvent(c->h2d, c->pev[1][p], 0), "slab wait");
        hip_check(hipMemcpyAsync(din[p], c->slabs.in[p], n * sl_in, hipMemcpyHostToDevice, c->h2d),
                  "upload slab");
        hip_check(hipEventRecord(c->pev[0][p], c->h2d), "event");
        hip_check(hipStreamWaitEvent(c->stream, c->pev[0][p], 0), "slab wait");
        if (k >= 2) hip_check(hipStreamWaitEvent(c->stream, c->pev[2][p], 0), "slab wait");
        c->ev = c->evs[p];
        c->evr = c->evrs[p];
        stage_fbp(c, din[p], n, na, bins, bins, filter, nullptr, dout[p]);
        hip_check(hipEventRecord(c->pev[1][p], c->stream), "event");
        hip_check(hipStreamWaitEvent(c->d2h, c->pev[1][p], 0), "slab wait");
        hip_check(hipMemcpyAsync(c->slabs.out[p], dout[p], n * sl_out, hipMemcpyDeviceToHost,
                                 c->d2h),
                  "download slab");
        hip_check(hipEventRecord(c->pev[2][p], c->d2h), "event");
    };
    auto finish = [&](int k) {
        const int p = k & 1, z0 = k * S, n = std::min(S, nz - z0);
        hip_check(hipEventSynchronize(c->pev[2][p]), "slab sync");
        c->ev = c->evs[p];
        c->evr = c->evrs[p];
        add_timings(c, false, true);
        cp->copy(out + (size_t)z0 * bins * bins, c->slabs.out[p], n * sl_out);
    };
    run_slabs(c, K, launch, finish);
}

}  // namespace cthost
