// capi.cpp — the C ABI of libct_hip.so (declared in include/ct_hip.h):
// context, HBM-budgeted device allocations, cached device state (trig tables,
// filter banks, twiddles, workspaces) and the stage orchestration of the
// reference's backend calls (IReconstructionBackend.h:10-55).
#include "ctx.h"

#include <cfloat>
#include <cmath>
#include <limits>
#include <cstdlib>
#include <cstdio>
#include <cstring>
#include <map>
#include <tuple>
#include <mutex>
#include <new>
#include <stdexcept>
#include <string>
#include <unordered_map>
#include <utility>
#include <vector>

using namespace cthip;
using namespace cthost;

namespace cthost {

void hip_check(hipError_t e, const char* what) {
    if (e != hipSuccess)
        throw StatusError(CT_ERR_HIP, std::string(what) + ": " + hipGetErrorString(e));
}

void* dmalloc(ct_hip_ctx* c, size_t bytes) {
    if (bytes == 0) bytes = 16;
    if (c->budget && c->in_use + bytes > c->budget)
        throw StatusError(CT_ERR_BUDGET, "HBM budget exceeded: need " + std::to_string(bytes) +
                                             " B, in use " + std::to_string(c->in_use) +
                                             " B of " + std::to_string(c->budget) + " B");
    void* p = nullptr;
    const hipError_t e = hipMalloc(&p, bytes);
    if (e != hipSuccess || !p) {
        (void)hipGetLastError();
        throw StatusError(CT_ERR_NOMEM, "hipMalloc(" + std::to_string(bytes) + ") failed");
    }
    c->sizes[p] = bytes;
    c->in_use += bytes;
    return p;
}

void dfree(ct_hip_ctx* c, void* p) {
    if (!p) return;
    auto it = c->sizes.find(p);
    if (it != c->sizes.end()) {
        c->in_use -= it->second;
        c->sizes.erase(it);
    }
    (void)hipFree(p);
}

void release(ct_hip_ctx* c, Buf& b) {
    dfree(c, b.p);
    b.p = nullptr;
    b.cap = 0;
}

// Serial's trig: th = degToRad(angle) in float, cos/sin by the C library's
// float functions (SerialBackend.cpp:92-95, 198-204; Utils.cpp:44-46). The
// table is built on the host because that is where those functions live; it
// then stays resident and is read through the scalar cache.
float deg2rad(float d) { return d * 3.14159265358979323846f / 180.0f; }

// At most kTrigCacheLists angle lists keep their tables on the device; the
// least recently used beyond that are freed at the start of the next call
// (trim_trig), never in the middle of one, so every table a call or a batch
// enqueues stays valid until the call returns. hipFree waits for the device,
// so a table still read by work in flight on another stream is freed after
// that work.
constexpr size_t kTrigCacheLists = 32;

void trim_trig(ct_hip_ctx* c) {
    while (c->trig.size() > kTrigCacheLists) {
        auto victim = c->trig.begin();
        for (auto it = c->trig.begin(); it != c->trig.end(); ++it)
            if (it->second.last_use < victim->second.last_use) victim = it;
        dfree(c, victim->second.cos);
        dfree(c, victim->second.sin);
        dfree(c, victim->second.alist);
        if (c->cur.cos == victim->second.cos) c->cur = ct_hip_ctx::Trig();
        c->trig.erase(victim);
    }
}

ct_hip_ctx::Trig get_trig(ct_hip_ctx* c, const float* angles_deg, int na) {
    std::vector<uint32_t> key(na);
    std::memcpy(key.data(), angles_deg, sizeof(float) * na);
    auto it = c->trig.find(key);
    if (it != c->trig.end()) {
        it->second.last_use = ++c->trig_tick;
        return it->second;
    }
    std::vector<float> cs(na), sn(na);
    for (int a = 0; a < na; ++a) {
        const float th = deg2rad(angles_deg[a]);
        cs[a] = std::cos(th);
        sn[a] = std::sin(th);
    }
    // The Radon stage's angle groups: |sin| <= |cos| (row walk over the
    // image), then the rest (row walk over the transposed image).
    std::vector<int> alist;
    alist.reserve(na);
    auto steep = [&](int a) { return std::fabs(sn[a]) > std::fabs(cs[a]); };
    for (int a = 0; a < na; ++a)
        if (!steep(a)) alist.push_back(a);
    const int n_row = (int)alist.size();
    for (int a = 0; a < na; ++a)
        if (steep(a)) alist.push_back(a);
    ct_hip_ctx::Trig t;
    t.cos = static_cast<float*>(dmalloc(c, sizeof(float) * na));
    t.sin = static_cast<float*>(dmalloc(c, sizeof(float) * na));
    t.alist = static_cast<int*>(dmalloc(c, sizeof(int) * na));
    t.n_row = n_row;
    t.last_use = ++c->trig_tick;
    hip_check(hipMemcpyAsync(t.cos, cs.data(), sizeof(float) * na, hipMemcpyHostToDevice, c->stream),
              "upload cos table");
    hip_check(hipMemcpyAsync(t.sin, sn.data(), sizeof(float) * na, hipMemcpyHostToDevice, c->stream),
              "upload sin table");
    hip_check(hipMemcpyAsync(t.alist, alist.data(), sizeof(int) * na, hipMemcpyHostToDevice,
                             c->stream),
              "upload angle list");
    hip_check(hipStreamSynchronize(c->stream), "trig upload");
    return c->trig[std::move(key)] = t;
}

// Sets the current call's table; trimming first keeps at most
// kTrigCacheLists tables on the device.
void set_trig(ct_hip_ctx* c, const float* angles_deg, int na) {
    trim_trig(c);
    c->cur = get_trig(c, angles_deg, na);
}

std::vector<float> serial_angles(int na) {
    std::vector<float> a(na);
    const float step = 180.0f / (float)na;  // SerialBackend.cpp:92-93, 127-130
    for (int i = 0; i < na; ++i) a[i] = step * (float)i;
    return a;
}

int norm_filter(int f) { return (f >= 0 && f <= 5) ? f : CT_FILTER_SHEPP_LOGAN; }

// Taps of the FFT row filter: the reference's createFilter taps with the
// Nyquist bin set to 0, which the filter then drops as Serial's inverse real
// FFT does (FFT.cpp:87-120 never reads bin n/2).
float* get_filter(ct_hip_ctx* c, int nfft, int type) {
    type = norm_filter(type);
    auto k = std::make_pair(nfft, type);
    auto it = c->filters.find(k);
    if (it != c->filters.end()) return it->second;
    std::vector<float> h(nfft / 2 + 1);
    serial_filter_taps(nfft, type, h.data());
    h[nfft / 2] = 0.0f;
    float* H = static_cast<float*>(dmalloc(c, sizeof(float) * h.size()));
    hip_check(hipMemcpyAsync(H, h.data(), sizeof(float) * h.size(), hipMemcpyHostToDevice,
                             c->stream),
              "upload filter taps");
    hip_check(hipStreamSynchronize(c->stream), "filter taps");
    c->filters[k] = H;
    return H;
}

float2* get_twiddles(ct_hip_ctx* c, int nfft) {
    auto it = c->twiddles.find(nfft);
    if (it != c->twiddles.end()) return it->second;
    float2* t = static_cast<float2*>(dmalloc(c, sizeof(float2) * (nfft - 1)));
    hip_check(launch_twiddles(nfft, t, c->stream), "twiddles");
    c->twiddles[nfft] = t;
    return t;
}

void wave_tables_host(const Geom& g, int filter, float2* tw, float* hs) {
    std::vector<float> taps(g.nfft / 2 + 1);
    serial_filter_taps(g.nfft, norm_filter(filter), taps.data());
    taps[g.nfft / 2] = 0.0f;
    for (int i = 0; i < kWaveHsFloats; ++i) hs[i] = 0.0f;
    wave_filter_tables(g.bins, g.sq, g.pad, g.nfft, taps.data(), tw, hs);
}

// Tables of the wave-form row filter, made on the host in FP64 from the
// reference's taps (FFT.cpp:87-120 semantics: Nyquist bin dropped).
static float2* get_wave_tables(ct_hip_ctx* c, const Geom& g, int type) {
    type = norm_filter(type);
    const auto k = std::make_pair(g.bins, type);
    auto it = c->wave_tabs.find(k);
    if (it != c->wave_tabs.end()) return it->second;
    std::vector<float2> tw(kWaveTabEntries);
    std::vector<float> hs(kWaveHsFloats);
    wave_tables_host(g, type, tw.data(), hs.data());
    const size_t tw_bytes = sizeof(float2) * kWaveTabEntries;
    float2* d = static_cast<float2*>(dmalloc(c, tw_bytes + sizeof(float) * kWaveHsFloats));
    hip_check(hipMemcpyAsync(d, tw.data(), tw_bytes, hipMemcpyHostToDevice, c->stream),
              "upload wave twiddles");
    hip_check(hipMemcpyAsync(reinterpret_cast<char*>(d) + tw_bytes, hs.data(),
                             sizeof(float) * kWaveHsFloats, hipMemcpyHostToDevice, c->stream),
              "upload wave filter");
    hip_check(hipStreamSynchronize(c->stream), "wave tables");
    c->wave_tabs[k] = d;
    return d;
}

int lds_filter_length(const Geom& g) {
    int L = 64;
    while (L < 2 * filter_kernel_reach(g.bins, g.sq, g.pad) + 1) L <<= 1;
    return std::min(L, g.nfft);
}

void lds_taps_host(const Geom& g, int filter, int L, float* H) {
    std::vector<float> taps(g.nfft / 2 + 1);
    serial_filter_taps(g.nfft, norm_filter(filter), taps.data());
    taps[g.nfft / 2] = 0.0f;
    if (L == g.nfft) std::copy(taps.begin(), taps.end(), H);
    else short_filter_taps(g.bins, g.sq, g.pad, g.nfft, taps.data(), L, H);
}

// Taps of the LDS FFT form at a length L shorter than Serial's nfft (see
// filter_kernel_reach), cached per (bins, type): bins in 23-26, 46-53 and
// 725-848 take them, six filter types, so the cache holds at most 816.
static float* get_short_taps(ct_hip_ctx* c, const Geom& g, int type, int L) {
    type = norm_filter(type);
    const auto k = std::make_pair(g.bins, type);
    auto it = c->short_taps.find(k);
    if (it != c->short_taps.end()) return it->second;
    std::vector<float> H(L / 2 + 1);
    lds_taps_host(g, type, L, H.data());
    float* d = static_cast<float*>(dmalloc(c, sizeof(float) * H.size()));
    hip_check(hipMemcpyAsync(d, H.data(), sizeof(float) * H.size(), hipMemcpyHostToDevice,
                             c->stream),
              "upload short taps");
    hip_check(hipStreamSynchronize(c->stream), "short taps");
    c->short_taps[k] = d;
    return d;
}

FiltArgs filter_args(ct_hip_ctx* c, const float* d_sino, int nz, int na, int bins, int filter,
                     float4* filt) {
    const Geom g = geometry(bins);
    const int r = wave_filter_r(bins, g.sq, g.pad, g.nfft);
    if (r > 0) {
        float2* t = get_wave_tables(c, g, filter);
        const float* hs = reinterpret_cast<const float*>(t + kWaveTabEntries);
        FiltArgs p = make_filt_args(d_sino, nz, na, bins, g.sq, g.pad, g.nfft, hs, t, filt);
        p.wave_r = r;
        return p;
    }
    check_fft(g);
    const int L = lds_filter_length(g);
    const float* H = L < g.nfft ? get_short_taps(c, g, filter, L) : get_filter(c, g.nfft, filter);
    return make_filt_args(d_sino, nz, na, bins, g.sq, g.pad, L, H, get_twiddles(c, L), filt);
}

hipError_t launch_filter(const FiltArgs& p, hipStream_t s) {
    if (p.wave_r > 0) return launch_row_filter_wave(p, s);
    return launch_row_filter_fft(p.sino, p.nz, p.na, p.bins, p.sq, p.pad, p.nfft, p.H, p.tw,
                                 p.filt, s);
}

// SerialBackend.cpp:152-161
Geom geometry(int bins) {
    Geom g;
    g.bins = bins;
    g.sq = (int)std::ceil(std::sqrt(2.0) * (double)bins);
    g.pad = g.sq / 2 - bins / 2;
    int p = 1;
    while (p < 2 * g.sq) p <<= 1;
    g.nfft = p > 64 ? p : 64;
    return g;
}

void check_fft(const Geom& g) {
    if (g.nfft > kMaxFFT)
        throw StatusError(CT_ERR_UNSUPPORTED,
                          "detector width " + std::to_string(g.bins) + " needs an FFT of " +
                              std::to_string(g.nfft) + " > " + std::to_string(kMaxFFT));
}

// Shape limits, checked before any device work. Radon: the walk's candidate
// intervals carry a margin (kDelta in radon.hip) proven against the float
// error of Serial's s for edges up to kRadonMaxEdge and bins up to
// kRadonMaxBins; larger shapes are refused rather than risk a lost deposit.
int shape_status(int stage, int w, int h, int bins, int backend, std::string* why) {
    auto fail = [&](int code, const std::string& m) {
        if (why) *why = m;
        return code;
    };
    if (stage == CT_STAGE_RADON) {
        if (w <= 0 || h <= 0 || bins <= 0) return fail(CT_ERR_INVALID, "Radon: empty shape");
        if (w > kRadonMaxEdge || h > kRadonMaxEdge || bins > kRadonMaxBins)
            return fail(CT_ERR_UNSUPPORTED,
                        "Radon: " + std::to_string(w) + " x " + std::to_string(h) + " px, " +
                            std::to_string(bins) + " bins is beyond the supported " +
                            std::to_string(kRadonMaxEdge) + " px edge / " +
                            std::to_string(kRadonMaxBins) + " bins");
        return CT_OK;
    }
    if (stage == CT_STAGE_FBP) {
        if (bins <= 0) return fail(CT_ERR_INVALID, "FBP: empty shape");
        if (backend == CT_FILTER_BACKEND_FFT_F32 && geometry(bins).nfft > kMaxFFT)
            return fail(CT_ERR_UNSUPPORTED, "detector width " + std::to_string(bins) +
                                                " needs an FFT longer than " +
                                                std::to_string(kMaxFFT));
        return CT_OK;
    }
    return fail(CT_ERR_INVALID, "unknown stage");
}

void check_radon(int w, int h, int bins) {
    std::string why;
    const int rc = shape_status(CT_STAGE_RADON, w, h, bins, 0, &why);
    if (rc != CT_OK) throw StatusError(rc, why);
}

void check_fbp(const ct_hip_ctx* c, int bins) {
    std::string why;
    const int rc = shape_status(CT_STAGE_FBP, 0, 0, bins, c->filter_backend, &why);
    if (rc != CT_OK) throw StatusError(rc, why);
}

size_t nzb_of(int nz) { return (size_t)((nz + kZB - 1) / kZB); }

// normalise + Radon of nz slices (w x h) into d_sino [nz][na][bins]; d_mm gets {min,max} per slice
struct RadonWs {
    unsigned* keys;
    float4 *row, *col;
    int* rflag;  // [zb][h][2] row extents of the row copy, then [zb][w][2] of the column copy
    float* side;
};

static RadonWs radon_ws(ct_hip_ctx* c, int w, int h, int nz, int na, int bins) {
    RadonWs ws;
    ws.keys = ensure<unsigned>(c, c->keys, sizeof(unsigned) * 2 * (size_t)nz);
    const size_t rzb = (size_t)((nz + kRZ - 1) / kRZ);
    ws.row = ensure<float4>(c, c->row, rzb * radon_zbytes(w, h));
    ws.col = ensure<float4>(c, c->col, rzb * radon_zbytes(h, w));
    ws.rflag = ensure<int>(c, c->rflag, sizeof(int) * 2 * rzb * ((size_t)h + w));
    const size_t nside = std::max(radon_side_floats(w, bins, nz, na),
                                  radon_side_floats(h, bins, nz, na));
    ws.side = nside ? ensure<float>(c, c->side, sizeof(float) * nside) : nullptr;
    return ws;
}

void stage_radon(ct_hip_ctx* c, const float* d_vol, int w, int h, int nz, int na, int bins,
                 float* d_sino, float* d_mm) {
    check_radon(w, h, bins);
    const RadonWs ws = radon_ws(c, w, h, nz, na, bins);
    unsigned* keys = ws.keys;
    float4 *row = ws.row, *col = ws.col;
    float* side = ws.side;
    hip_check(hipEventRecord(c->evr[0], c->stream), "event");
    hip_check(launch_minmax_init(keys, nz, c->stream), "minmax init");
    hip_check(launch_minmax(d_vol, (size_t)w * h, nz, keys, c->stream), "minmax");
    hip_check(launch_minmax_decode(keys, nz, d_mm, c->stream), "minmax decode");
    const int nzb = (nz + kRZ - 1) / kRZ;
    int* rflag = ws.rflag;                         // rows of the row copy
    int* cflag = ws.rflag + (size_t)nzb * h * 2;  // rows of the column copy (image columns)
    hip_check(launch_zero_ints(rflag, 2 * nzb * (h + w), c->stream), "zero row extents");
    hip_check(hipEventRecord(c->evr[1], c->stream), "event");
    hip_check(launch_normalise(d_vol, w, h, nz, d_mm, row, col, rflag, cflag, c->stream),
              "normalise");
    // The two walks write disjoint angles of the sinogram. They run one after
    // the other: side by side on two streams they evict each other's cache
    // lines (measured 24% slower at 1024^3, no gain at 512^3).
    hip_check(hipEventRecord(c->evr[2], c->stream), "event");
    const int n_row = c->cur.n_row;
    if (radon_walks_paired(w, h)) {
        // both walks in one launch: its time is the image walk's, the
        // transpose's is zero
        hip_check(launch_radon_walks(row, col, rflag, cflag, w, h, nz, na, bins, c->cur.cos,
                                     c->cur.sin, c->cur.alist, n_row, d_sino, side, c->stream),
                  "radon walks");
        hip_check(hipEventRecord(c->evr[3], c->stream), "event");
        hip_check(hipEventRecord(c->evr[4], c->stream), "event");
        c->radon_timed = true;
        return;
    }
    hip_check(launch_radon_rows(row, rflag, w, h, nz, na, bins, c->cur.cos, c->cur.sin,
                                c->cur.alist, n_row, d_sino, side, c->stream),
              "radon rows");
    hip_check(hipEventRecord(c->evr[3], c->stream), "event");
    hip_check(launch_radon_rows(col, cflag, h, w, nz, na, bins, c->cur.sin, c->cur.cos,
                                c->cur.alist + n_row, na - n_row, d_sino, side, c->stream),
              "radon rows of the transpose");
    hip_check(hipEventRecord(c->evr[4], c->stream), "event");
    c->radon_timed = true;
}

const ct_hip_ctx::Toep& get_toeplitz(ct_hip_ctx* c, const Geom& g, int type) {
    type = norm_filter(type);
    const auto key = std::make_tuple(g.nfft, type, g.bins);
    auto it = c->toeplitz.find(key);
    if (it != c->toeplitz.end()) return it->second;
    ct_hip_ctx::Toep t;
    t.Np = (g.sq + 255) / 256 * 256;
    t.Kp = (g.bins + 1 + 31) / 32 * 32;  // dp has bins+1 entries
    const size_t n = (size_t)t.Np * t.Kp;
    t.hi = static_cast<unsigned short*>(dmalloc(c, 2 * n));
    t.lo = static_cast<unsigned short*>(dmalloc(c, 2 * n));
    t.f8 = static_cast<unsigned char*>(dmalloc(c, n + 8));
    t.tmax = static_cast<float*>(dmalloc(c, sizeof(float)));
    double* H = static_cast<double*>(dmalloc(c, sizeof(double) * (g.nfft / 2 + 1)));
    double* h = static_cast<double*>(dmalloc(c, sizeof(double) * 2 * g.nfft));  // h, then G
    float* T32 = static_cast<float*>(dmalloc(c, sizeof(float) * n));
    hip_check(launch_toeplitz_design(g.nfft, type, g.sq, g.bins, g.pad, t.Np, t.Kp, H, h, t.hi,
                                     t.lo, T32, t.f8, t.tmax, c->stream),
              "toeplitz design");
    hip_check(hipStreamSynchronize(c->stream), "sync");
    dfree(c, H);
    dfree(c, h);
    dfree(c, T32);
    return c->toeplitz[key] = t;
}

void launch_gemm_filter(ct_hip_ctx* c, const float* d_sino, int nz, int na, int bins, int filter,
                        float4* filt, float* rmax, hipStream_t stream) {
    const Geom g = geometry(bins);
    const ct_hip_ctx::Toep& t = get_toeplitz(c, g, filter);
    if (c->filter_backend == CT_FILTER_BACKEND_GEMM_FP8) {
        need(rmax != nullptr, "FP8 filter: no row-maximum buffer");
        hip_check(launch_row_absmax(d_sino, (size_t)nz * na, bins, rmax, stream), "row absmax");
    }
    hip_check(launch_toeplitz_gemm(c->filter_backend, d_sino, nz, na, bins, g.sq, t.Np, t.Kp, t.hi,
                                   t.lo, t.f8, rmax, t.tmax, filt, stream),
              "toeplitz gemm");
}

// row filter of nz sinograms [nz][na][bins] -> filt [nz/4][na][sq][4]
void stage_filter(ct_hip_ctx* c, const float* d_sino, int nz, int na, int bins, int filter,
                  float4* filt) {
    if (c->filter_backend == CT_FILTER_BACKEND_FFT_F32) {
        hip_check(launch_filter(filter_args(c, d_sino, nz, na, bins, filter, filt), c->stream),
                  "row filter");
        return;
    }
    float* rmax = c->filter_backend == CT_FILTER_BACKEND_GEMM_FP8
                      ? ensure<float>(c, c->rmax, sizeof(float) * (size_t)nz * na)
                      : nullptr;
    launch_gemm_filter(c, d_sino, nz, na, bins, filter, filt, rmax, c->stream);
}

// Backprojection tile configuration for (N, na): the fixed one, or the
// autotuner's (which times candidates on the context stream the first time a
// class is seen, so callers ask before their first timing event).
int bp_config_for(ct_hip_ctx* c, int N, int na) {
    return c->bp_config_fixed ? c->bp_config : autotune(c, N, na, nullptr);
}

// filter + backprojection of nz sinograms [nz][na][bins] onto N x N
static float4* fbp_ws(ct_hip_ctx* c, int nz, int na, int bins) {
    if (c->filter_backend == CT_FILTER_BACKEND_GEMM_FP8)
        (void)ensure<float>(c, c->rmax, sizeof(float) * (size_t)nz * na);
    return ensure<float4>(c, c->filt,
                          sizeof(float4) * nzb_of(nz) * (size_t)na * geometry(bins).sq);
}

void reserve_stages(ct_hip_ctx* c, int w, int h, int nz, int na, int bins, bool with_radon,
                    bool with_fbp) {
    if (with_radon) (void)radon_ws(c, w, h, nz, na, bins);
    if (with_fbp) (void)fbp_ws(c, nz, na, bins);
}

void stage_fbp(ct_hip_ctx* c, const float* d_sino, int nz, int na, int bins, int N, int filter,
               const float* d_mm, float* d_out) {
    const Geom g = geometry(bins);
    check_fbp(c, bins);
    const int cfg = bp_config_for(c, N, na);
    float4* filt = fbp_ws(c, nz, na, bins);
    hip_check(hipEventRecord(c->ev[1], c->stream), "event");
    stage_filter(c, d_sino, nz, na, bins, filter, filt);
    hip_check(hipEventRecord(c->ev[2], c->stream), "event");
    const float scale = 3.14159265358979323846f / (2.0f * (float)na);  // SerialBackend.cpp:246
    hip_check(launch_backproject(filt, nz, na, g.sq, N, c->cur.cos, c->cur.sin, scale, d_mm, d_out, c->stream,
                                 cfg),
              "backprojection");
    hip_check(hipEventRecord(c->ev[3], c->stream), "event");
}

void zero_timings(ct_hip_ctx* c) {
    for (double& t : c->timings) t = 0;
    for (double& t : c->radon_ms) t = 0;
    c->radon_slabs = 0;
}

// Adds the stage times recorded in the current event set: the Radon stage
// (ev[0] -> ev[1], its walks in evr) when with_radon, the filter (ev[1] ->
// ev[2]) and the backprojection (ev[2] -> ev[3]) when with_fbp.
void add_timings(ct_hip_ctx* c, bool with_radon, bool with_fbp) {
    add_timings(c, c->ev, c->evr, with_radon, with_fbp);
}

void add_timings(ct_hip_ctx* c, const hipEvent_t* ev, const hipEvent_t* evr, bool with_radon,
                 bool with_fbp) {
    hip_check(hipEventSynchronize(with_fbp ? ev[3] : ev[1]), "event sync");
    float t01 = 0, t12 = 0, t23 = 0;
    if (with_radon) {
        ++c->radon_slabs;
        hip_check(hipEventElapsedTime(&t01, ev[0], ev[1]), "elapsed");
        for (int k = 0; k < 4; ++k) {
            float r = 0;
            hip_check(hipEventElapsedTime(&r, evr[k], evr[k + 1]), "elapsed");
            c->radon_ms[k] += r;
        }
    }
    if (with_fbp) {
        hip_check(hipEventElapsedTime(&t12, ev[1], ev[2]), "elapsed");
        hip_check(hipEventElapsedTime(&t23, ev[2], ev[3]), "elapsed");
    }
    c->timings[0] += t01;
    c->timings[1] += t12;
    c->timings[2] += t23;
    c->timings[3] += (double)t01 + t12 + t23;
}

// Whole-buffer host transfers through the pinned stager, ordered after the
// work already enqueued on the context's stream; both return when done.
void to_device(ct_hip_ctx* c, void* d_dst, const void* h_src, size_t bytes) {
    hip_check(hipEventRecord(c->mark, c->stream), "event");
    stager(c).upload(d_dst, h_src, bytes, c->mark);
}

void to_host(ct_hip_ctx* c, void* h_dst, const void* d_src, size_t bytes) {
    hip_check(hipEventRecord(c->mark, c->stream), "event");
    stager(c).download(h_dst, d_src, bytes, c->mark);
}

void bind_device(ct_hip_ctx* c) { hip_check(hipSetDevice(c->device), "hipSetDevice"); }

void need(bool ok, const char* msg) {
    if (!ok) throw StatusError(CT_ERR_INVALID, msg);
}

// ---- phantom parameters (project spec) -----------------------------------
uint64_t splitmix64(uint64_t x) {
    x += 0x9E3779B97F4A7C15ull;
    x = (x ^ (x >> 30)) * 0xBF58476D1CE4E5B9ull;
    x = (x ^ (x >> 27)) * 0x94D049BB133111EBull;
    return x ^ (x >> 31);
}

float uniform_draw(uint64_t seed, int e, int p, double lo, double hi) {
    const uint64_t x = splitmix64(seed ^ splitmix64((uint64_t)e * 16u + (uint64_t)p));
    const double u = (double)(x >> 11) * (1.0 / 9007199254740992.0);
    return (float)(lo + (hi - lo) * u);
}

PhEll derive(const ct_ellipsoid& e) {
    const double a = e.phi, b = e.theta, g = e.psi;
    const double ca = std::cos(a), sa = std::sin(a), cb = std::cos(b), sb = std::sin(b),
                 cg = std::cos(g), sg = std::sin(g);
    const double Rz1[3][3] = {{ca, -sa, 0}, {sa, ca, 0}, {0, 0, 1}};
    const double Ry[3][3] = {{cb, 0, sb}, {0, 1, 0}, {-sb, 0, cb}};
    const double Rz2[3][3] = {{cg, -sg, 0}, {sg, cg, 0}, {0, 0, 1}};
    double T[3][3], R[3][3];
    for (int i = 0; i < 3; ++i)
        for (int j = 0; j < 3; ++j) {
            double s = 0.0;
            for (int k = 0; k < 3; ++k) s += Rz1[i][k] * Ry[k][j];
            T[i][j] = s;
        }
    for (int i = 0; i < 3; ++i)
        for (int j = 0; j < 3; ++j) {
            double s = 0.0;
            for (int k = 0; k < 3; ++k) s += T[i][k] * Rz2[k][j];
            R[i][j] = s;
        }
    PhEll q;
    q.cx = e.cx;
    q.cy = e.cy;
    q.cz = e.cz;
    // body coordinates q = R^T (p - c)
    for (int i = 0; i < 3; ++i)
        for (int j = 0; j < 3; ++j) q.m[i * 3 + j] = (float)R[j][i];
    const double ax[3] = {e.ax, e.ay, e.az};
    for (int k = 0; k < 3; ++k) q.ia[k] = (float)(1.0 / (ax[k] * ax[k]));
    q.rho = e.rho;
    const double cc[3] = {e.cx, e.cy, e.cz};
    for (int j = 0; j < 3; ++j) {
        double s = 0.0;
        for (int k = 0; k < 3; ++k) s += (R[j][k] * ax[k]) * (R[j][k] * ax[k]);
        const double ext = std::sqrt(s) + 1e-3;
        q.lo[j] = (float)(cc[j] - ext);
        q.hi[j] = (float)(cc[j] + ext);
    }
    return q;
}

void phantom_to_device(ct_hip_ctx* c, const ct_ellipsoid* e, int n, int nx, int ny, int nz,
                       float* d_out) {
    need(nx > 0 && ny > 0 && nz > 0 && n >= 0 && (n == 0 || e), "phantom: bad arguments");
    std::vector<PhEll> h(n > 0 ? n : 1);
    for (int k = 0; k < n; ++k) h[k] = derive(e[k]);
    PhEll* d = ensure<PhEll>(c, c->ell, sizeof(PhEll) * h.size());
    hip_check(hipMemcpyAsync(d, h.data(), sizeof(PhEll) * h.size(), hipMemcpyHostToDevice,
                             c->stream),
              "upload ellipsoids");
    hip_check(launch_phantom(d, n, nx, ny, nz, d_out, c->stream), "phantom");
    hip_check(hipStreamSynchronize(c->stream), "phantom sync");
}

}  // namespace

// =========================================================================
extern "C" {

int ct_hip_device_count(void) {
    int n = 0;
    if (hipGetDeviceCount(&n) != hipSuccess) {
        (void)hipGetLastError();
        return 0;
    }
    return n;
}

// The context's events that order work without timing it.
static std::vector<hipEvent_t*> sync_events(ct_hip_ctx* c) {
    std::vector<hipEvent_t*> v{&c->mark};
    for (int p = 0; p < 2; ++p) {
        v.push_back(&c->up_done[p]);
        v.push_back(&c->comp_done[p]);
    }
    for (int t = 0; t < 3; ++t) {
        v.push_back(&c->dl_done[t]);
        for (auto& e : c->dl_piece[t]) v.push_back(&e);
    }
    return v;
}

int ct_hip_create(int device, size_t hbm_budget_bytes, ct_hip_ctx** out) {
    if (!out) return CT_ERR_INVALID;
    *out = nullptr;
    const int n = ct_hip_device_count();
    if (n <= 0) return CT_ERR_NODEVICE;
    if (device < 0 || device >= n) return CT_ERR_INVALID;
    ct_hip_ctx* c = new (std::nothrow) ct_hip_ctx();
    if (!c) return CT_ERR_NOMEM;
    c->device = device;
    c->budget = hbm_budget_bytes;
    if (hipSetDevice(device) != hipSuccess || hipStreamCreateWithFlags(&c->stream, hipStreamNonBlocking) != hipSuccess) {
        delete c;
        return CT_ERR_HIP;
    }
    cthost::noalloc_init(device);
    bool ok = hipStreamCreateWithFlags(&c->h2d, hipStreamNonBlocking) == hipSuccess &&
              hipStreamCreateWithFlags(&c->d2h, hipStreamNonBlocking) == hipSuccess;
    for (auto& set : c->evs)
        for (auto& e : set) ok = ok && hipEventCreate(&e) == hipSuccess;
    for (auto& set : c->evrs)
        for (auto& e : set) ok = ok && hipEventCreate(&e) == hipSuccess;
    for (hipEvent_t* e : sync_events(c))
        ok = ok && hipEventCreateWithFlags(e, hipEventDisableTiming) == hipSuccess;
    if (!ok) {
        ct_hip_destroy(c);
        return CT_ERR_HIP;
    }
    *out = c;
    return CT_OK;
}

void ct_hip_destroy(ct_hip_ctx* c) {
    if (!c) return;
    (void)hipSetDevice(c->device);
    if (c->stream) (void)hipStreamSynchronize(c->stream);
    try {
        cthost::batch_release(c);
    } catch (...) {
    }
    if (c->h2d) (void)hipStreamSynchronize(c->h2d);
    if (c->d2h) (void)hipStreamSynchronize(c->d2h);
    try {
        cthost::release_host_paths(c);
    } catch (...) {
    }
    for (auto& kv : c->sizes) (void)hipFree(kv.first);
    c->sizes.clear();
    for (auto& set : c->evs)
        for (auto& e : set)
            if (e) (void)hipEventDestroy(e);
    for (auto& set : c->evrs)
        for (auto& e : set)
            if (e) (void)hipEventDestroy(e);
    for (hipEvent_t* e : sync_events(c))
        if (*e) (void)hipEventDestroy(*e);
    if (c->h2d) (void)hipStreamDestroy(c->h2d);
    if (c->d2h) (void)hipStreamDestroy(c->d2h);
    if (c->stream) (void)hipStreamDestroy(c->stream);
    delete c;
}

const char* ct_hip_last_error(const ct_hip_ctx* c) { return c ? c->err.c_str() : "null context"; }

size_t ct_hip_bytes_in_use(const ct_hip_ctx* c) { return c ? c->in_use : 0; }

int ct_hip_clear_workspace(ct_hip_ctx* c) {
    return guarded(c, [&] {
        hip_check(hipStreamSynchronize(c->stream), "sync");
        for (Buf* b : {&c->vol, &c->row, &c->col, &c->rflag, &c->side, &c->sino, &c->filt, &c->out,
                       &c->keys, &c->mm, &c->ell, &c->pc_cnt, &c->pc_pts, &c->pc_coord, &c->met})
            release(c, *b);
        for (auto& kv : c->trig) {
            dfree(c, kv.second.cos);
            dfree(c, kv.second.sin);
            dfree(c, kv.second.alist);
        }
        c->trig.clear();
        c->cur = ct_hip_ctx::Trig();
        for (auto& kv : c->filters) dfree(c, kv.second);
        for (auto& kv : c->toeplitz) {
            dfree(c, kv.second.hi);
            dfree(c, kv.second.lo);
            dfree(c, kv.second.f8);
            dfree(c, kv.second.tmax);
        }
        c->toeplitz.clear();
        release(c, c->rmax);
        batch_release(c);
        release_host_paths(c);
        for (auto& kv : c->twiddles) dfree(c, kv.second);
        for (auto& kv : c->wave_tabs) dfree(c, kv.second);
        for (auto& kv : c->short_taps) dfree(c, kv.second);
        c->filters.clear();
        c->twiddles.clear();
        c->wave_tabs.clear();
        c->short_taps.clear();
    });
}

int ct_hip_alloc(ct_hip_ctx* c, size_t bytes, void** dptr) {
    return guarded(c, [&] {
        need(dptr != nullptr, "alloc: null out pointer");
        *dptr = dmalloc(c, bytes);
    });
}

int ct_hip_free(ct_hip_ctx* c, void* p) {
    return guarded(c, [&] {
        need(c->sizes.count(p) > 0, "free: pointer not owned by this context");
        hip_check(hipStreamSynchronize(c->stream), "sync");
        dfree(c, p);
    });
}

// Both go through the context's pinned stager, ordered after the work
// already enqueued on the context's stream.
int ct_hip_upload(ct_hip_ctx* c, void* dst, const void* src, size_t bytes) {
    return guarded(c, [&] {
        need(dst && src, "upload: null pointer");
        hip_check(hipEventRecord(c->mark, c->stream), "event");
        stager(c).upload(dst, src, bytes, c->mark);
    });
}

int ct_hip_download(ct_hip_ctx* c, void* dst, const void* src, size_t bytes) {
    return guarded(c, [&] {
        need(dst && src, "download: null pointer");
        hip_check(hipEventRecord(c->mark, c->stream), "event");
        stager(c).download(dst, src, bytes, c->mark);
    });
}

int ct_hip_synchronize(ct_hip_ctx* c) {
    return guarded(c, [&] { hip_check(hipStreamSynchronize(c->stream), "sync"); });
}

int ct_phantom_params(uint64_t seed, int n, ct_ellipsoid* out) {
    if (n < 0 || (n > 0 && !out)) return CT_ERR_INVALID;
    const double pi = 3.14159265358979323846;
    for (int e = 0; e < n; ++e) {
        ct_ellipsoid& q = out[e];
        q.cx = uniform_draw(seed, e, 0, -0.6, 0.6);
        q.cy = uniform_draw(seed, e, 1, -0.6, 0.6);
        q.cz = uniform_draw(seed, e, 2, -0.6, 0.6);
        q.ax = uniform_draw(seed, e, 3, 0.05, 0.4);
        q.ay = uniform_draw(seed, e, 4, 0.05, 0.4);
        q.az = uniform_draw(seed, e, 5, 0.05, 0.4);
        q.phi = uniform_draw(seed, e, 6, 0.0, pi);
        q.theta = uniform_draw(seed, e, 7, 0.0, pi);
        q.psi = uniform_draw(seed, e, 8, 0.0, pi);
        q.rho = uniform_draw(seed, e, 9, -0.3, 1.0);
    }
    return CT_OK;
}

int ct_hip_phantom_dev(ct_hip_ctx* c, const ct_ellipsoid* e, int n, int nx, int ny, int nz,
                       float* d_out) {
    return guarded(c, [&] {
        need(d_out != nullptr, "phantom: null output");
        phantom_to_device(c, e, n, nx, ny, nz, d_out);
    });
}

int ct_hip_phantom(ct_hip_ctx* c, const ct_ellipsoid* e, int n, int nx, int ny, int nz,
                   float* host_out) {
    return guarded(c, [&] {
        need(host_out != nullptr, "phantom: null output");
        need(nx > 0 && ny > 0 && nz > 0, "phantom: bad shape");
        const size_t bytes = sizeof(float) * (size_t)nx * ny * nz;
        float* d = ensure<float>(c, c->vol, bytes);
        phantom_to_device(c, e, n, nx, ny, nz, d);
        to_host(c, host_out, d, bytes);
    });
}

int ct_hip_noisy_log_dev(ct_hip_ctx* c, float* d_sino, size_t n, float I0, float kappa,
                         uint64_t seed) {
    return guarded(c, [&] {
        need(d_sino && n > 0 && I0 > 0.0f && kappa > 0.0f, "noisyLog: bad arguments");
        hip_check(launch_noisy_log(d_sino, n, I0, kappa, seed, c->stream), "noisy log");
        hip_check(hipStreamSynchronize(c->stream), "sync");
    });
}

int ct_hip_radon_dev(ct_hip_ctx* c, const float* d_vol, int w, int h, int d, int na, int bins,
                     float* d_sino, float* minmax_out) {
    return guarded(c, [&] {
        need(d_vol && d_sino, "radon: null pointer");
        need(w > 0 && h > 0 && d > 0 && na > 0 && bins > 0, "radon: empty input");
        check_radon(w, h, bins);
        const std::vector<float> ang = serial_angles(na);
        set_trig(c, ang.data(), na);
        float* d_mm = ensure<float>(c, c->mm, sizeof(float) * 2 * (size_t)d);
        stage_radon(c, d_vol, w, h, d, na, bins, d_sino, d_mm);
        if (minmax_out)
            hip_check(hipMemcpyAsync(minmax_out, d_mm, sizeof(float) * 2 * (size_t)d,
                                     hipMemcpyDeviceToHost, c->stream),
                      "download min/max");
        hip_check(hipStreamSynchronize(c->stream), "sync");
    });
}

int ct_hip_compute_sinogram(ct_hip_ctx* c, const float* slice, int w, int h, int na, int bins,
                            float* sino_out, float* minmax_out) {
    return guarded(c, [&] {
        need(slice && sino_out, "computeSinogram: null pointer");
        need(w > 0 && h > 0 && na > 0 && bins > 0, "computeSinogram: empty input");
        check_radon(w, h, bins);
        const size_t pix = (size_t)w * h;
        float* d_in = ensure<float>(c, c->vol, sizeof(float) * pix);
        float* d_sino = ensure<float>(c, c->sino, sizeof(float) * (size_t)na * bins);
        float* d_mm = ensure<float>(c, c->mm, sizeof(float) * 2);
        const std::vector<float> ang = serial_angles(na);
        set_trig(c, ang.data(), na);
        hip_check(hipMemcpyAsync(d_in, slice, sizeof(float) * pix, hipMemcpyHostToDevice, c->stream),
                  "upload slice");
        stage_radon(c, d_in, w, h, 1, na, bins, d_sino, d_mm);
        hip_check(hipMemcpyAsync(sino_out, d_sino, sizeof(float) * (size_t)na * bins,
                                 hipMemcpyDeviceToHost, c->stream),
                  "download sinogram");
        float mm[2];
        hip_check(hipMemcpyAsync(mm, d_mm, sizeof(mm), hipMemcpyDeviceToHost, c->stream), "mm");
        hip_check(hipStreamSynchronize(c->stream), "sync");
        if (minmax_out) {
            minmax_out[0] = mm[0];
            minmax_out[1] = mm[1];
        }
    });
}

int ct_hip_reconstruct_slice(ct_hip_ctx* c, const float* sino, int na, int bins,
                             const float* angles_deg, int N, int filter, float* out) {
    return guarded(c, [&] {
        need(sino && angles_deg && out, "reconstructSlice: null pointer");
        need(na > 0 && bins > 0 && N > 0, "reconstructSlice: empty input");
        check_fbp(c, bins);
        const size_t n_in = (size_t)na * bins;
        float* d_sino = ensure<float>(c, c->sino, sizeof(float) * n_in);
        float* d_out = ensure<float>(c, c->out, sizeof(float) * (size_t)N * N);
        set_trig(c, angles_deg, na);
        hip_check(hipMemcpyAsync(d_sino, sino, sizeof(float) * n_in, hipMemcpyHostToDevice,
                                 c->stream),
                  "upload sinogram");
        stage_fbp(c, d_sino, 1, na, bins, N, filter, nullptr, d_out);
        hip_check(hipMemcpyAsync(out, d_out, sizeof(float) * (size_t)N * N, hipMemcpyDeviceToHost,
                                 c->stream),
                  "download slice");
        hip_check(hipStreamSynchronize(c->stream), "sync");
    });
}

// The shapes a volume call will run, checked before any allocation or launch.
static void check_volume(ct_hip_ctx* c, int w, int h, int d, int na) {
    need(w > 0 && h > 0 && d > 0 && na > 0, "reconstructVolume: empty input");
    check_radon(w, h, w);  // bins = w, SerialBackend.cpp:287-288
    if (h == w) check_fbp(c, w);
}

// Slices per slab of the device path: the whole volume when its workspace
// fits the context's budget, else the most slices (a multiple of 8, so the
// Radon z-blocks and the filter's row pairs are those of the whole-volume
// run and the result is the same bit for bit) whose workspace does.
static int volume_dev_slab(ct_hip_ctx* c, int w, int h, int d, int na, bool sino_ws) {
    if (!c->budget) return d;
    const size_t per = volume_ws_per_slice(w, h, na, sino_ws);
    size_t held = 0;  // workspaces the call reuses count as available
    for (const Buf* b : {&c->row, &c->col, &c->filt, &c->sino}) held += b->cap;
    const size_t avail = c->budget > c->in_use - held ? c->budget - (c->in_use - held) : 0;
    const size_t fit = avail / std::max<size_t>(per, 1) * 15 / 16;  // margin for the small buffers
    if (fit >= (size_t)d) return d;
    if (fit < (size_t)std::min(d, kRZ))
        throw StatusError(CT_ERR_BUDGET, "reconstructVolume: the HBM budget holds no slab of " +
                                             std::to_string(std::min(d, kRZ)) + " slices");
    return (int)(fit / kRZ * kRZ);
}

static void volume_dev(ct_hip_ctx* c, const float* d_vol, int w, int h, int d, int na, int filter,
                       float* d_out, float* d_sino_out) {
    need(d_vol && d_out, "reconstructVolume: null pointer");
    check_volume(c, w, h, d, na);
    const int bins = w;  // SerialBackend.cpp:287-288
    const std::vector<float> ang = serial_angles(na);
    set_trig(c, ang.data(), na);
    if (h == w) (void)bp_config_for(c, w, na);
    zero_timings(c);
    const int S = volume_dev_slab(c, w, h, d, na, d_sino_out == nullptr);
    const size_t wh = (size_t)w * h, sl_sino = (size_t)na * bins;
    float* d_mm = ensure<float>(c, c->mm, sizeof(float) * 2 * (size_t)d);
    float* ws_sino = d_sino_out ? nullptr : ensure<float>(c, c->sino, sizeof(float) * S * sl_sino);
    for (int z0 = 0; z0 < d; z0 += S) {
        const int nz = std::min(S, d - z0);
        float* sino = d_sino_out ? d_sino_out + (size_t)z0 * sl_sino : ws_sino;
        hip_check(hipEventRecord(c->ev[0], c->stream), "event");
        stage_radon(c, d_vol + z0 * wh, w, h, nz, na, bins, sino, d_mm + 2 * (size_t)z0);
        if (h == w) {
            stage_fbp(c, sino, nz, na, bins, w, filter, d_mm + 2 * (size_t)z0, d_out + z0 * wh);
            add_timings(c, true, true);
        } else {
            // Buffer3D::setSlice drops slices whose shape differs (CTTypes.h:112-116);
            // Serial still spends the sinogram time, which is what is recorded
            hip_check(hipEventRecord(c->ev[1], c->stream), "event");
            hip_check(hipMemsetAsync(d_out + z0 * wh, 0, sizeof(float) * nz * wh, c->stream),
                      "zero");
            add_timings(c, true, false);
        }
    }
    hip_check(hipStreamSynchronize(c->stream), "sync");
}

int ct_hip_reconstruct_volume_dev(ct_hip_ctx* c, const float* d_vol, int w, int h, int d, int na,
                                  int filter, float* d_out, float* d_sino_out) {
    return guarded(c, [&] {
        volume_dev(c, d_vol, w, h, d, na, filter, d_out, d_sino_out);
        hip_check(hipStreamSynchronize(c->stream), "sync");
    });
}

int ct_hip_reconstruct_volume(ct_hip_ctx* c, const float* vol, int w, int h, int d, int na,
                              int filter, float* out, float* sino_out) {
    return guarded(c, [&] {
        need(vol && out, "reconstructVolume: null pointer");
        check_volume(c, w, h, d, na);
        volume_host(c, vol, w, h, d, na, filter, out, sino_out);
    });
}

static void from_sinos_dev(ct_hip_ctx* c, const float* d_sinos, int nz, int na, int bins,
                           const float* angles_deg, int filter, float* d_out) {
    need(d_sinos && angles_deg && d_out, "reconstructVolumeFromSinograms: null pointer");
    need(nz > 0 && na > 0 && bins > 0, "reconstructVolumeFromSinograms: empty input");
    check_fbp(c, bins);
    set_trig(c, angles_deg, na);
    (void)bp_config_for(c, bins, na);
    zero_timings(c);
    hip_check(hipEventRecord(c->ev[1], c->stream), "event");
    stage_fbp(c, d_sinos, nz, na, bins, bins, filter, nullptr, d_out);
    add_timings(c, false, true);
}

int ct_hip_reconstruct_from_sinograms_dev(ct_hip_ctx* c, const float* d_sinos, int nz, int na,
                                          int bins, const float* angles_deg, int filter,
                                          float* d_out) {
    return guarded(c, [&] { from_sinos_dev(c, d_sinos, nz, na, bins, angles_deg, filter, d_out); });
}

int ct_hip_reconstruct_from_sinograms(ct_hip_ctx* c, const float* sinos, int nz, int na, int bins,
                                      const float* angles_deg, int filter, float* out) {
    return guarded(c, [&] {
        need(sinos && angles_deg && out, "reconstructVolumeFromSinograms: null pointer");
        need(nz > 0 && na > 0 && bins > 0, "reconstructVolumeFromSinograms: empty input");
        check_fbp(c, bins);
        from_sinos_host(c, sinos, nz, na, bins, angles_deg, filter, out);
    });
}

static void point_cloud_dev(ct_hip_ctx* c, const float* d_vol, int w, int h, int d,
                            const float* xc, const float* yc, const float* zc, float thr,
                            float* out, size_t cap, size_t* count) {
    need(d_vol && count, "extractPointCloud: null pointer");
    need(w > 0 && h > 0 && d > 0, "extractPointCloud: empty input");
    const size_t n = (size_t)w * h * d;
    need(n < (1ull << 32), "extractPointCloud: volume too large");
    const unsigned nb = pc_blocks(n);
    unsigned* cnt = ensure<unsigned>(c, c->pc_cnt, sizeof(unsigned) * 2 * (size_t)nb);
    hip_check(launch_pc_count(d_vol, n, thr, cnt, c->stream), "point count");
    std::vector<unsigned> hc(nb);
    hip_check(hipMemcpyAsync(hc.data(), cnt, sizeof(unsigned) * nb, hipMemcpyDeviceToHost,
                             c->stream),
              "download counts");
    hip_check(hipStreamSynchronize(c->stream), "sync");
    size_t total = 0;
    for (unsigned b = 0; b < nb; ++b) {
        const unsigned k = hc[b];
        hc[b] = (unsigned)total;
        total += k;
    }
    *count = total;
    if (!out || cap < total || total == 0) return;
    unsigned* offs = cnt + nb;
    hip_check(hipMemcpyAsync(offs, hc.data(), sizeof(unsigned) * nb, hipMemcpyHostToDevice,
                             c->stream),
              "upload offsets");
    float* coords = nullptr;
    const float* dx = nullptr;
    const float* dy = nullptr;
    const float* dz = nullptr;
    if (xc || yc || zc) {
        coords = ensure<float>(c, c->pc_coord, sizeof(float) * ((size_t)w + h + d));
        if (xc) {
            hip_check(hipMemcpyAsync(coords, xc, sizeof(float) * w, hipMemcpyHostToDevice, c->stream), "xc");
            dx = coords;
        }
        if (yc) {
            hip_check(hipMemcpyAsync(coords + w, yc, sizeof(float) * h, hipMemcpyHostToDevice, c->stream), "yc");
            dy = coords + w;
        }
        if (zc) {
            hip_check(hipMemcpyAsync(coords + w + h, zc, sizeof(float) * d, hipMemcpyHostToDevice, c->stream), "zc");
            dz = coords + w + h;
        }
    }
    float4* pts = ensure<float4>(c, c->pc_pts, sizeof(float4) * total);
    hip_check(launch_pc_write(d_vol, n, w, h, thr, offs, dx, dy, dz, pts, c->stream), "point write");
    hip_check(hipMemcpyAsync(out, pts, sizeof(float4) * total, hipMemcpyDeviceToHost, c->stream),
              "download points");
    hip_check(hipStreamSynchronize(c->stream), "sync");
}

int ct_hip_extract_point_cloud_dev(ct_hip_ctx* c, const float* d_vol, int w, int h, int d,
                                   const float* xc, const float* yc, const float* zc,
                                   float threshold, float* out, size_t cap, size_t* count) {
    return guarded(c, [&] { point_cloud_dev(c, d_vol, w, h, d, xc, yc, zc, threshold, out, cap, count); });
}

int ct_hip_extract_point_cloud(ct_hip_ctx* c, const float* vol, int w, int h, int d,
                               const float* xc, const float* yc, const float* zc, float threshold,
                               float* out, size_t cap, size_t* count) {
    return guarded(c, [&] {
        need(vol != nullptr && count != nullptr, "extractPointCloud: null pointer");
        need(w > 0 && h > 0 && d > 0, "extractPointCloud: empty input");
        const size_t n = (size_t)w * h * d;
        float* d_vol = ensure<float>(c, c->vol, sizeof(float) * n);
        to_device(c, d_vol, vol, sizeof(float) * n);
        point_cloud_dev(c, d_vol, w, h, d, xc, yc, zc, threshold, out, cap, count);
    });
}

// CtReconstructionController.cpp:1199-1277: the sums come from the device
// pass; the means, PSNR and the max-difference rule are the controller's
// double arithmetic, in its order.
static void metrics_dev(ct_hip_ctx* c, const float* d_orig, const float* d_recon, size_t n,
                        ct_metrics* out) {
    need(out != nullptr, "volumeMetrics: null output");
    MetricsSums m{0.0, 0.0, 0.0f, FLT_MAX, -FLT_MAX, 0.0f};
    if (n > 0) {
        need(d_orig && d_recon, "volumeMetrics: null pointer");
        need(((uintptr_t)d_orig & 15) == 0 && ((uintptr_t)d_recon & 15) == 0,
             "volumeMetrics: device pointers must be 16-byte aligned");
        MetricsSums* part =
            ensure<MetricsSums>(c, c->met, sizeof(MetricsSums) * (kMetricsBlocks + 1));
        hip_check(launch_metrics(d_orig, d_recon, n, part, part + kMetricsBlocks, c->stream),
                  "metrics");
        hip_check(hipMemcpyAsync(&m, part + kMetricsBlocks, sizeof(m), hipMemcpyDeviceToHost,
                                 c->stream),
                  "download metrics");
        hip_check(hipStreamSynchronize(c->stream), "sync");
    }
    float max_abs = m.mx;
    if (max_abs < 1e-6f) max_abs = 1.0f;
    const double mse = n > 0 ? m.s2 / (double)n : 0.0;
    const double mae = n > 0 ? m.s1 / (double)n : 0.0;
    const double range = (double)(m.hi - m.lo);
    out->mse = mse;
    out->mae = mae;
    out->psnr = (mse <= 0.0 || range <= 0.0) ? std::numeric_limits<double>::infinity()
                                             : 10.0 * std::log10((range * range) / mse);
    out->max_abs = max_abs;
    out->vmin = m.lo;
    out->vmax = m.hi;
    out->count = (uint64_t)n;
}

int ct_hip_volume_metrics_dev(ct_hip_ctx* c, const float* d_orig, const float* d_recon, size_t n,
                              ct_metrics* out) {
    return guarded(c, [&] { metrics_dev(c, d_orig, d_recon, n, out); });
}

int ct_hip_volume_metrics(ct_hip_ctx* c, const float* orig, const float* recon, size_t n,
                          ct_metrics* out) {
    return guarded(c, [&] {
        if (n == 0) return metrics_dev(c, nullptr, nullptr, 0, out);
        need(orig && recon, "volumeMetrics: null pointer");
        float* d_o = ensure<float>(c, c->vol, sizeof(float) * n);
        float* d_r = ensure<float>(c, c->out, sizeof(float) * n);
        to_device(c, d_o, orig, sizeof(float) * n);
        to_device(c, d_r, recon, sizeof(float) * n);
        metrics_dev(c, d_o, d_r, n, out);
    });
}

// CtReconstructionController.cpp:1353-1482 (exportAllSinograms): the
// sinogram of every slice with num_angles angles and bins = w, written as
// (d, num_angles, w) with the sidecar; the sidecar's min/max are the extremes
// of the slices' original min/max.
int ct_hip_export_sinograms(ct_hip_ctx* c, const float* vol, int w, int h, int d, int na,
                            const char* npy_path) {
    return guarded(c, [&] {
        need(vol && npy_path, "exportSinograms: null pointer");
        need(w > 0 && h > 0 && d > 0 && na > 0, "exportSinograms: empty input");
        check_radon(w, h, w);
        const size_t pix = (size_t)w * h * d, ns = (size_t)d * na * w;
        float* d_vol = ensure<float>(c, c->vol, sizeof(float) * pix);
        float* d_sino = ensure<float>(c, c->sino, sizeof(float) * ns);
        float* d_mm = ensure<float>(c, c->mm, sizeof(float) * 2 * (size_t)d);
        const std::vector<float> ang = serial_angles(na);
        set_trig(c, ang.data(), na);
        to_device(c, d_vol, vol, sizeof(float) * pix);
        stage_radon(c, d_vol, w, h, d, na, w, d_sino, d_mm);
        std::vector<float> sino(ns), mm(2 * (size_t)d);
        to_host(c, sino.data(), d_sino, sizeof(float) * ns);
        hip_check(hipMemcpyAsync(mm.data(), d_mm, sizeof(float) * mm.size(), hipMemcpyDeviceToHost,
                                 c->stream),
                  "download min/max");
        hip_check(hipStreamSynchronize(c->stream), "sync");
        float lo = FLT_MAX, hi = -FLT_MAX;
        for (int z = 0; z < d; ++z) {
            if (mm[2 * z] < lo) lo = mm[2 * z];
            if (mm[2 * z + 1] > hi) hi = mm[2 * z + 1];
        }
        write_sinogram_files(npy_path, sino.data(), d, na, w, lo, hi);
    });
}

int ct_hip_design_filter(ct_hip_ctx* c, int n, int filter, float* out) {
    return guarded(c, [&] {
        need(out != nullptr, "designFilter: null output");
        need(n >= 4 && (n & (n - 1)) == 0, "designFilter: n must be a power of two >= 4");
        serial_filter_taps(n, norm_filter(filter), out);
    });
}

int ct_hip_set_filter_backend(ct_hip_ctx* c, int backend) {
    return guarded(c, [&] {
        need(backend >= CT_FILTER_BACKEND_FFT_F32 && backend <= CT_FILTER_BACKEND_GEMM_FP8,
             "setFilterBackend: unknown backend");
        c->filter_backend = backend;
    });
}

int ct_hip_filter_sinograms(ct_hip_ctx* c, const float* sinos, int nz, int na, int bins,
                            int filter, float* out) {
    return guarded(c, [&] {
        need(sinos && out, "filterSinograms: null pointer");
        need(nz > 0 && na > 0 && bins > 0, "filterSinograms: empty input");
        check_fbp(c, bins);
        const Geom g = geometry(bins);
        const size_t nin = (size_t)nz * na * bins;
        const size_t nf = nzb_of(nz) * (size_t)na * g.sq;
        float* d_in = ensure<float>(c, c->sino, sizeof(float) * nin);
        float4* filt = ensure<float4>(c, c->filt, sizeof(float4) * nf);
        to_device(c, d_in, sinos, sizeof(float) * nin);
        stage_filter(c, d_in, nz, na, bins, filter, filt);
        std::vector<float> tmp(4 * nf);
        to_host(c, tmp.data(), filt, sizeof(float4) * nf);
        hip_check(hipStreamSynchronize(c->stream), "sync");
        for (int z = 0; z < nz; ++z)
            for (int a = 0; a < na; ++a)
                for (int i = 0; i < g.sq; ++i)
                    out[((size_t)z * na + a) * g.sq + i] =
                        tmp[(((size_t)(z / kZB) * na + a) * g.sq + i) * kZB + (z % kZB)];
    });
}

int ct_hip_last_timings(const ct_hip_ctx* c, double* ms4) {
    if (!c || !ms4) return CT_ERR_INVALID;
    for (int i = 0; i < 4; ++i) ms4[i] = c->timings[i];
    return CT_OK;
}

int ct_hip_check_shape(int stage, int w, int h, int bins, int filter_backend) {
    return shape_status(stage, w, h, bins, filter_backend, nullptr);
}

int ct_hip_last_radon_timings(const ct_hip_ctx* c, double* ms2, int* launches) {
    if (!c || !ms2) return CT_ERR_INVALID;
    if (!c->radon_timed) return CT_ERR_INVALID;
    ms2[0] = c->radon_ms[2];
    ms2[1] = c->radon_ms[3];
    if (launches) *launches = c->radon_slabs;
    return CT_OK;
}

int ct_hip_last_kernel_timings(const ct_hip_ctx* c, double* ms, int n) {
    if (!c || !ms || n < 0) return CT_ERR_INVALID;
    const double v[6] = {c->radon_ms[0], c->radon_ms[1], c->radon_ms[2], c->radon_ms[3],
                         c->timings[1], c->timings[2]};
    for (int i = 0; i < n && i < 6; ++i) ms[i] = v[i];
    return CT_OK;
}

}  // extern "C"
