// batch.cpp — batched form of reconstructVolumeFromSinograms
// (IReconstructionBackend.h:36-41): many volumes of uneven size (edge, depth,
// angle count, filter) reconstructed by two persistent launches, the row
// filter over all (volume, angle, z-block) items and the backprojection over
// all (volume, z-block, tile) items, both fed from device work lists through
// atomic counters. The pair (plus the counter reset) is captured once into a
// hipGraph and replayed for later batches: only the work lists change, and
// they are uploaded before the replay, outside the graph.
//
// The backprojection tile configuration (voxel tile, angles per LDS stage)
// comes from the autotuner, which times every instantiated configuration on
// a synthetic volume of the job's class and caches the winner per (N, na):
// in the context, for every context of the process on the same device (a
// tuning takes up to a second at 1024^2 x 1440 angles, and a server's
// context, made with the server, would otherwise tune each class again while
// its jobs wait), and optionally in a file named by CT_HIP_TUNE_CACHE.
#include "ctx.h"

#include <algorithm>
#include <cstdio>
#include <cstdlib>
#include <fstream>
#include <mutex>
#include <numeric>
#include <sstream>
#include <tuple>

using namespace cthip;

namespace cthost {

// Per-stream batch launcher state. The work-list buffers are sized for
// kMaxBatch jobs once, so captured graphs (keyed by the launch shape) stay
// valid across batches; the tables are staged through pinned host memory.
BatchSlot::~BatchSlot() = default;

void slot_init(ct_hip_ctx* c, BatchSlot* s, bool own_stream) {
    if (s->ready) return;
    if (own_stream) {
        hip_check(hipStreamCreateWithFlags(&s->stream, hipStreamNonBlocking), "slot stream");
        s->own_stream = true;
    } else {
        s->stream = c->stream;
    }
    hip_check(hipEventCreate(&s->ev_start), "slot event");
    hip_check(hipEventCreate(&s->ev_done), "slot event");
    s->ftab = static_cast<FiltJob*>(dmalloc(c, sizeof(FiltJob) * kMaxBatch));
    s->btab = static_cast<BpJob*>(dmalloc(c, sizeof(BpJob) * kMaxBatch));
    s->hdr = static_cast<int*>(dmalloc(c, sizeof(int) * (2 + 3 * kMaxFftClasses)));
    s->counters = static_cast<int*>(dmalloc(c, sizeof(int) * (kBpCounters + kMaxFftClasses)));
    hip_check(hipHostMalloc((void**)&s->h_ftab, sizeof(FiltJob) * kMaxBatch, hipHostMallocDefault),
              "pinned table");
    hip_check(hipHostMalloc((void**)&s->h_btab, sizeof(BpJob) * kMaxBatch, hipHostMallocDefault),
              "pinned table");
    hip_check(hipHostMalloc((void**)&s->h_hdr, sizeof(int) * (2 + 3 * kMaxFftClasses),
                            hipHostMallocDefault),
              "pinned hdr");
    s->ready = true;
}

void slot_release(ct_hip_ctx* c, BatchSlot* s) {
    if (!s->ready) return;
    if (s->stream) (void)hipStreamSynchronize(s->stream);
    for (auto& kv : s->graphs) (void)hipGraphExecDestroy(kv.second);
    s->graphs.clear();
    dfree(c, s->ftab);
    dfree(c, s->btab);
    dfree(c, s->hdr);
    dfree(c, s->counters);
    release(c, s->rmax);
    (void)hipHostFree(s->h_ftab);
    (void)hipHostFree(s->h_btab);
    (void)hipHostFree(s->h_hdr);
    (void)hipEventDestroy(s->ev_start);
    (void)hipEventDestroy(s->ev_done);
    if (s->own_stream) (void)hipStreamDestroy(s->stream);
    *s = BatchSlot();
}

struct BatchGraph {
    BatchSlot slot;
    Buf filt;
};

void batch_release(ct_hip_ctx* c) {
    if (!c->batch) return;
    slot_release(c, &c->batch->slot);
    release(c, c->batch->filt);
    delete c->batch;
    c->batch = nullptr;
}

// Tuning results of the process: (device, N, na) -> configuration. Every
// configuration gives the same result bit for bit, so sharing them changes
// speed only.
static std::mutex g_tune_mu;
static std::map<std::tuple<int, int, int>, int> g_tuned;

static void share_tuned(int device, int N, int na, int cfg) {
    std::lock_guard<std::mutex> lk(g_tune_mu);
    g_tuned[std::make_tuple(device, N, na)] = cfg;
}

static bool shared_tuned(int device, int N, int na, int* cfg) {
    std::lock_guard<std::mutex> lk(g_tune_mu);
    auto it = g_tuned.find(std::make_tuple(device, N, na));
    if (it == g_tuned.end()) return false;
    *cfg = it->second;
    return true;
}

static void load_tune_cache(ct_hip_ctx* c) {
    if (c->tune_cache_loaded) return;
    c->tune_cache_loaded = true;
    const char* path = std::getenv("CT_HIP_TUNE_CACHE");
    if (!path) return;
    std::ifstream f(path);
    int N, na, cfg;
    while (f >> N >> na >> cfg)
        if (cfg >= 0 && cfg < kNumBpConfigs) c->tuned[{N, na}] = cfg;
}

static void store_tune_cache(int N, int na, int cfg) {
    const char* path = std::getenv("CT_HIP_TUNE_CACHE");
    if (!path) return;
    std::ofstream f(path, std::ios::app);
    f << N << " " << na << " " << cfg << "\n";
}

// Times every configuration on a synthetic volume of N x N, na angles, deep
// enough for 131072 tiles of 16 x 16 voxels x 4 slices, within 1 GiB and
// 6.4e10 voxel-updates (336 slices at 512^2 x 720, 40 at 1024^2 x 1440, where
// a tuning takes about a second). Shallower volumes rank by their last, partly
// filled wave of workgroups: over 128 slices at 512^2 x 720 the 8-slice
// 32-voxel tiles fill 2.7 waves and time 4.6% slower than the 4-slice
// ones, while over the full 512 slices they are 5.5% faster.
int autotune(ct_hip_ctx* c, int N, int na, std::vector<double>* times) {
    load_tune_cache(c);
    auto it = c->tuned.find({N, na});
    if (it != c->tuned.end() && !times) return it->second;
    int shared = 0;
    if (!times && shared_tuned(c->device, N, na, &shared)) return c->tuned[{N, na}] = shared;
    const Geom g = geometry(N);
    const int tiles16 = ((N + 15) / 16) * ((N + 15) / 16);
    int nzb = std::max(1, 8192 * 16 / std::max(1, tiles16));
    const size_t per_zb = (size_t)na * g.sq * sizeof(float4) + sizeof(float) * kZB * (size_t)N * N;
    nzb = (int)std::max<size_t>(1, std::min<size_t>((size_t)nzb, (size_t)(1u << 30) / per_zb));
    nzb = std::max(1, std::min(nzb, (int)(6.4e10 / ((double)N * N * na * kZB))));
    const int nz = nzb * kZB;
    float4* filt = nullptr;
    float* out = nullptr;
    try {
        filt = static_cast<float4*>(dmalloc(c, (size_t)nzb * na * g.sq * sizeof(float4)));
        out = static_cast<float*>(dmalloc(c, sizeof(float) * (size_t)nz * N * N));
    } catch (const StatusError& e) {
        // a budgeted context without room to tune keeps the default tiles
        dfree(c, filt);
        if (e.code != CT_ERR_BUDGET && e.code != CT_ERR_NOMEM) throw;
        return c->bp_config;
    }
    std::vector<double> ms(kNumBpConfigs, 1e30);
    try {
        hip_check(hipMemsetAsync(filt, 0, (size_t)nzb * na * g.sq * sizeof(float4), c->stream), "zero");
        std::vector<float> ang(na);
        for (int a = 0; a < na; ++a) ang[a] = (180.0f / (float)na) * (float)a;
        const ct_hip_ctx::Trig t = get_trig(c, ang.data(), na);
        BpArgs p;
        p.filt = filt;
        p.out = out;
        p.cos_t = t.cos;
        p.sin_t = t.sin;
        p.mm = nullptr;
        p.nz = nz;
        p.na = na;
        p.sq = g.sq;
        p.N = N;
        p.scale = 1.0f;
        p.check = bp_needs_check(N, g.sq) ? 1 : 0;
        hipEvent_t e0 = nullptr, e1 = nullptr;
        hip_check(hipEventCreate(&e0), "event");
        if (hipEventCreate(&e1) != hipSuccess) {
            (void)hipEventDestroy(e0);
            hip_check(hipErrorOutOfMemory, "event");
        }
        try {
            for (int cfg = 0; cfg < kNumBpConfigs; ++cfg) {
                hip_check(launch_backproject_cfg(cfg, p, c->stream), "autotune warmup");
                for (int r = 0; r < 3; ++r) {
                    hip_check(hipEventRecord(e0, c->stream), "event");
                    hip_check(launch_backproject_cfg(cfg, p, c->stream), "autotune");
                    hip_check(hipEventRecord(e1, c->stream), "event");
                    hip_check(hipEventSynchronize(e1), "event sync");
                    float e = 0;
                    hip_check(hipEventElapsedTime(&e, e0, e1), "elapsed");
                    ms[cfg] = std::min(ms[cfg], (double)e);
                }
            }
        } catch (...) {
            (void)hipEventDestroy(e0);
            (void)hipEventDestroy(e1);
            throw;
        }
        (void)hipEventDestroy(e0);
        (void)hipEventDestroy(e1);
    } catch (...) {
        dfree(c, filt);
        dfree(c, out);
        throw;
    }
    dfree(c, filt);
    dfree(c, out);
    const int best = (int)(std::min_element(ms.begin(), ms.end()) - ms.begin());
    c->tuned[{N, na}] = best;
    share_tuned(c->device, N, na, best);
    store_tune_cache(N, na, best);
    if (times) *times = ms;
    return best;
}

size_t filt_elems(const ct_fbp_job& j) {
    const Geom g = geometry(j.bins);
    return (size_t)((j.nz + kZB - 1) / kZB) * j.na * g.sq;
}

size_t job_filt_bytes(const ct_fbp_job& j) { return sizeof(float4) * filt_elems(j); }

// Builds the work lists for jobs[0..nj) (filtered sinograms go to filts[i])
// and enqueues, on the slot's stream, the row filter (persistent launches,
// one per transform class, for the FFT backend; one Toeplitz-GEMM launch per
// job for the MFMA backends) and the persistent backprojection, with the
// counter reset and the persistent launches as a graph replay when flags ask
// for it. Does not wait. The caller holds c->mu.
void enqueue_batch(ct_hip_ctx* c, BatchSlot* s, const ct_fbp_job* jobs, int nj, float4* const* filts,
                   int flags) {
    need(jobs != nullptr && nj > 0, "fbpBatch: no jobs");
    need(nj <= kMaxBatch, "fbpBatch: too many jobs in one batch");
    trim_trig(c);  // before this batch's tables are looked up, never between
    // The context's row filter backend: the FFT forms run as persistent
    // launches over the batch's rows; a Toeplitz-GEMM form (MFMA) runs one
    // launch per job, ahead of the captured backprojection.
    const bool gemm = c->filter_backend != CT_FILTER_BACKEND_FFT_F32;
    long long work_best = -1;
    size_t max_rows = 0;
    int dom = 0;
    for (int i = 0; i < nj; ++i) {
        const ct_fbp_job& j = jobs[i];
        need(j.d_sinos && j.d_out && j.angles_deg, "fbpBatch: null pointer in job");
        need(j.nz > 0 && j.na > 0 && j.bins > 0, "fbpBatch: empty job");
        check_fbp(c, j.bins);
        max_rows = std::max(max_rows, (size_t)j.nz * j.na);
        const long long work = (long long)j.bins * j.bins * j.nz * j.na;
        if (work > work_best) {
            work_best = work;
            dom = i;
        }
    }
    const int cfg = (flags & CT_BATCH_AUTOTUNE) ? autotune(c, jobs[dom].bins, jobs[dom].na, nullptr)
                                                : c->bp_config;
    const int T = kBpConfigs[cfg].tile, Z2 = kBpConfigs[cfg].zb2;
    std::vector<int> tiles(nj);
    std::vector<BpJob> bj(nj);
    std::vector<FiltJob> fj(nj);
    for (int i = 0; i < nj; ++i) {
        const ct_fbp_job& j = jobs[i];
        const Geom g = geometry(j.bins);
        const ct_hip_ctx::Trig t = get_trig(c, j.angles_deg, j.na);
        if (gemm) (void)get_toeplitz(c, g, j.filter);  // made on the context stream
        else fj[i].a = filter_args(c, j.d_sinos, j.nz, j.na, j.bins, j.filter, filts[i]);
        BpArgs& p = bj[i].a;
        p.filt = filts[i];
        p.out = j.d_out;
        p.cos_t = t.cos;
        p.sin_t = t.sin;
        p.mm = nullptr;
        p.nz = j.nz;
        p.na = j.na;
        p.sq = g.sq;
        p.N = j.bins;
        p.scale = 3.14159265358979323846f / (2.0f * (float)j.na);
        p.check = bp_needs_check(j.bins, g.sq) ? 1 : 0;
        tiles[i] = ((j.bins + T - 1) / T) * ((j.bins + T - 1) / T);
    }
    // filter taps / twiddles / trig tables are made on the context stream
    if (s->stream != c->stream) hip_check(hipStreamSynchronize(c->stream), "sync device state");
    // costlier items (more angles) first, so the tail of the launch is short
    std::vector<int> order(nj);
    std::iota(order.begin(), order.end(), 0);
    std::stable_sort(order.begin(), order.end(), [&](int x, int y) { return jobs[x].na > jobs[y].na; });
    int bitems = 0;
    for (int k = 0; k < nj; ++k) {
        const int i = order[k];
        s->h_btab[k] = bj[i];
        s->h_btab[k].item_begin = bitems;
        bitems += (((jobs[i].nz + kZB - 1) / kZB + Z2 - 1) / Z2) * tiles[i];
    }
    // Filter work lists, one per transform class (LDS FFT length, or wave
    // form R): a persistent launch reserves LDS and registers for its class,
    // so classes are not mixed in one launch.
    std::map<int, std::vector<int>> classes;  // filter_class -> jobs
    if (!gemm)
        for (int i = 0; i < nj; ++i) classes[filter_class(fj[i].a)].push_back(i);
    need((int)classes.size() <= kMaxFftClasses, "fbpBatch: too many row filter classes");
    std::vector<int> cls_key, cls_first;
    int pos = 0, ci = 0;
    for (auto& kv : classes) {
        int items = 0;
        cls_key.push_back(kv.first);
        cls_first.push_back(pos);
        for (int i : kv.second) {
            s->h_ftab[pos] = fj[i];
            s->h_ftab[pos].item_begin = items;
            items += ((jobs[i].nz + kZB - 1) / kZB) * jobs[i].na;
            ++pos;
        }
        s->h_hdr[2 + 3 * ci] = cls_first.back();
        s->h_hdr[3 + 3 * ci] = (int)kv.second.size();
        s->h_hdr[4 + 3 * ci] = items;
        ++ci;
    }
    s->h_hdr[0] = nj;
    s->h_hdr[1] = bitems;
    if (!gemm)
        hip_check(hipMemcpyAsync(s->ftab, s->h_ftab, sizeof(FiltJob) * nj, hipMemcpyHostToDevice,
                                 s->stream),
                  "upload filter work list");
    hip_check(hipMemcpyAsync(s->btab, s->h_btab, sizeof(BpJob) * nj, hipMemcpyHostToDevice,
                             s->stream),
              "upload backprojection work list");
    hip_check(hipMemcpyAsync(s->hdr, s->h_hdr, sizeof(int) * (2 + 3 * ci), hipMemcpyHostToDevice,
                             s->stream),
              "upload header");
    std::vector<int> fgrid(ci);
    for (int k = 0; k < ci; ++k) {
        fgrid[k] = filter_persistent_grid(cls_key[k], c->device);
        need(fgrid[k] > 0, "fbpBatch: no occupancy for the persistent row filter");
    }
    const int bgrid = bp_persistent_grid(cfg, c->device);
    need(bgrid > 0, "fbpBatch: no occupancy for the persistent backprojection");
    auto enqueue = [&]() {
        hip_check(launch_zero_ints(s->counters, kBpCounters + ci, s->stream), "zero counters");
        // kernel arguments depend only on the graph key (class count, FFT
        // lengths, grids); where each class's jobs start is in the header
        for (int k = 0; k < ci; ++k)
            hip_check(launch_row_filter_persistent(s->ftab, s->hdr + 2 + 3 * k,
                                                   s->counters + kBpCounters + k, cls_key[k], fgrid[k],
                                                   s->stream),
                      "persistent row filter");
        hip_check(launch_bp_persistent(cfg, s->btab, s->hdr, s->counters, bgrid, s->stream),
                  "persistent backprojection");
    };
    hip_check(hipEventRecord(s->ev_start, s->stream), "event");
    if (gemm) {
        // one buffer of row maxima for the FP8 form, sized for the largest
        // job before any launch (jobs use it one after another on the stream)
        float* rmax = c->filter_backend == CT_FILTER_BACKEND_GEMM_FP8
                          ? ensure<float>(c, s->rmax, sizeof(float) * max_rows)
                          : nullptr;
        for (int i = 0; i < nj; ++i) {
            const ct_fbp_job& j = jobs[i];
            launch_gemm_filter(c, j.d_sinos, j.nz, j.na, j.bins, j.filter, filts[i], rmax, s->stream);
        }
    }
    if (flags & CT_BATCH_GRAPH) {
        std::vector<int> key{cfg, bgrid, gemm ? 1 : 0};
        for (int k = 0; k < ci; ++k) {
            key.push_back(cls_key[k]);
            key.push_back(fgrid[k]);
        }
        auto it = s->graphs.find(key);
        if (it == s->graphs.end()) {
            hipGraph_t graph = nullptr;
            hip_check(hipStreamBeginCapture(s->stream, hipStreamCaptureModeRelaxed), "begin capture");
            try {
                enqueue();
            } catch (...) {
                (void)hipStreamEndCapture(s->stream, &graph);
                if (graph) (void)hipGraphDestroy(graph);
                throw;
            }
            hip_check(hipStreamEndCapture(s->stream, &graph), "end capture");
            hipGraphExec_t exec = nullptr;
            const hipError_t e = hipGraphInstantiate(&exec, graph, nullptr, nullptr, 0);
            (void)hipGraphDestroy(graph);
            hip_check(e, "instantiate");
            it = s->graphs.emplace(key, exec).first;
        }
        hip_check(hipGraphLaunch(it->second, s->stream), "graph launch");
    } else {
        enqueue();
    }
    hip_check(hipEventRecord(s->ev_done, s->stream), "event");
}

static void run_batch(ct_hip_ctx* c, const ct_fbp_job* jobs, int nj, int flags) {
    need(jobs != nullptr && nj > 0, "fbpBatch: no jobs");
    if (!c->batch) c->batch = new BatchGraph();
    BatchGraph* b = c->batch;
    slot_init(c, &b->slot, false);
    size_t total = 0;
    std::vector<size_t> off(nj);
    for (int i = 0; i < nj; ++i) {
        need(jobs[i].nz > 0 && jobs[i].na > 0 && jobs[i].bins > 0, "fbpBatch: empty job");
        off[i] = total;
        total += filt_elems(jobs[i]);
    }
    float4* filt = ensure<float4>(c, b->filt, sizeof(float4) * std::max<size_t>(total, 1));
    std::vector<float4*> f(nj);
    for (int i = 0; i < nj; ++i) f[i] = filt + off[i];
    enqueue_batch(c, &b->slot, jobs, nj, f.data(), flags);
    hip_check(hipEventSynchronize(b->slot.ev_done), "event sync");
    float t = 0;
    hip_check(hipEventElapsedTime(&t, b->slot.ev_start, b->slot.ev_done), "elapsed");
    c->timings[0] = 0;
    c->timings[1] = 0;
    c->timings[2] = 0;
    c->timings[3] = t;
}

}  // namespace cthost

using namespace cthost;

extern "C" {

int ct_hip_fbp_batch_dev(ct_hip_ctx* c, const ct_fbp_job* jobs, int njobs, int flags) {
    return guarded(c, [&] { run_batch(c, jobs, njobs, flags); });
}

int ct_hip_autotune(ct_hip_ctx* c, int N, int na, int* config_out, double* ms_out, int n_ms) {
    return guarded(c, [&] {
        need(N > 0 && na > 0, "autotune: bad shape");
        std::vector<double> t;
        const int best = autotune(c, N, na, ms_out ? &t : nullptr);
        if (config_out) *config_out = best;
        if (ms_out)
            for (int i = 0; i < n_ms && i < (int)t.size(); ++i) ms_out[i] = t[i];
    });
}

int ct_hip_bp_config_count(void) { return kNumBpConfigs; }

int ct_hip_bp_config(int i, int* tile, int* angles, int* slices) {
    if (i < 0 || i >= kNumBpConfigs || !tile || !angles) return CT_ERR_INVALID;
    *tile = kBpConfigs[i].tile;
    *angles = kBpConfigs[i].angles;
    if (slices) *slices = kZB * kBpConfigs[i].zb2;
    return CT_OK;
}

int ct_hip_set_bp_config(ct_hip_ctx* c, int cfg) {
    return guarded(c, [&] {
        need(cfg >= 0 && cfg < kNumBpConfigs, "setBpConfig: unknown configuration");
        c->bp_config = cfg;
        c->bp_config_fixed = true;
    });
}

}  // extern "C"
