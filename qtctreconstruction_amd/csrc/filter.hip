// filter.hip — row filter of the FBP: the padded FFT convolution of every
// sinogram row, and the twiddle table.
//
// Reference semantics (SerialBackend.cpp:150-195, FilterDesign.cpp:13-111,
// FFT.cpp:51-120): each row of `bins` samples is placed at offset
// pad = sq/2 - bins/2 in a zero row of nfft = max(64, nextPow2(2*sq)),
// transformed, multiplied by the real filter H (n/2+1 taps, H[0] = 0),
// transformed back, and the first sq samples are kept. The reference's
// inverse never reads the Nyquist bin, so the effective filter also has
// H[n/2] = 0. H holds the reference's own taps (FP32 design, computed on the
// host once per (n, type)) with that bin zeroed.
//
// Filtering: the filter is real and even, so two real rows ride in one
// complex FFT (row A in the real part, row B in the imaginary part) and come
// back separated. A block owns one angle of one z-block (4 rows -> 2 complex
// FFTs) held in LDS: decimation-in-frequency forward from natural-order input
// (bit-reversed spectrum), multiply by H at bit-reversed positions,
// decimation-in-time inverse back to natural order, so rows are loaded and
// stored contiguously and no element is moved to its bit-reversed place.
// Radix-2 stages run fused three at a time in registers. Twiddles are read
// from a stage-ordered table in global memory (small enough to stay in L1),
// at consecutive addresses across a wave.
// Output is written z-innermost (float4 = 4 slices per bin), the layout the
// backprojection stages from.
#include "ct_internal.h"

#include <atomic>

namespace cthip {

// Twiddles in stage order: the table of a butterfly span m (m = 2, 4, ..., n)
// starts at m/2 - 1 and holds W_m^j = exp(-2 pi i j / m) for j < m/2, so a
// stage reads its twiddles at consecutive addresses. W_m^j equals
// W_n^(j n/m) bit for bit: the FP64 angles differ by exact power-of-two
// scalings only.
__global__ void twiddle_kernel(int n, float2* __restrict__ tw) {
    const int idx = blockIdx.x * blockDim.x + threadIdx.x;
    if (idx >= n - 1) return;
    const int lg = 31 - __clz(idx + 1);  // floor(log2(idx + 1))
    const int m = 2 << lg;
    const int j = idx + 1 - m / 2;
    const double a = -2.0 * 3.14159265358979323846 * (double)j / (double)m;
    tw[idx] = make_float2((float)cos(a), (float)sin(a));
}

// Complex products with one rounded product and one FMA per component
// (4 VALU instead of 6).
__device__ __forceinline__ float2 cmul(float2 a, float2 b) {
    return make_float2(fmaf(a.x, b.x, -(a.y * b.y)), fmaf(a.x, b.y, a.y * b.x));
}
__device__ __forceinline__ float2 cmulc(float2 a, float2 b) {  // a * conj(b)
    return make_float2(fmaf(a.x, b.x, a.y * b.y), fmaf(a.y, b.x, -(a.x * b.y)));
}

// Global-memory views of the argument pointers: the persistent kernel reads
// its arguments from a work list in memory, where the compiler only sees
// generic pointers; naming the address space keeps its loads and stores
// global (not flat), as in the per-launch kernel.
typedef float f2v __attribute__((ext_vector_type(2)));
typedef float f4v __attribute__((ext_vector_type(4)));
typedef const float __attribute__((address_space(1))) g_cfloat;
typedef const f2v __attribute__((address_space(1))) g_cfloat2;
typedef f4v __attribute__((address_space(1))) g_float4;

// LDS slot of transform element e: bits 0-4 of e are XORed with bits 3-7, a
// permutation within each 32-element (256-byte) block. The butterfly passes
// read elements base + r*h; for small spans h, the lanes of one b64 read
// would otherwise land on a few of the 64 banks (up to 8-way conflicts at
// h = 1, 4); swizzled, the mean conflict degree over every pass of n = 64 ..
// 4096 is 1.07 (from 4.1).
__device__ __forceinline__ int swz(int e) { return e ^ ((e >> 3) & 31); }

// Radix-2 butterfly stages on the two LDS transforms X0, X1 of N points.
// Stage of half-span m, element e at position q = e mod 2m < m paired with
// e + m; W = W_2m^q (stage order: W_M^j at M/2 - 1 + j), conjugated when
// CONJ (the inverse transform):
//   DIT (increasing spans, bit-reversed in, natural out):
//       t = W*b;  a' = a + t;  b' = a - t
//   DIF (decreasing spans, natural in, bit-reversed out):
//       a' = a + b;  b' = W*(a - b)
// A fused pass applies 3 consecutive stages (half-spans H, 2H, 4H) to groups
// of 8 elements in registers, then writes them back; every stage's
// operations and their order are those of the single stage. N and H are
// compile-time constants, so the group and LDS-slot arithmetic folds into
// shifts, masks and immediate offsets. When log2 N is not a multiple of 3,
// the 1 or 2 innermost stages run on groups of 2 or 4 at half-spans 1 and 2,
// whose twiddles are exactly 1 and -i (applied as such).
template <bool CONJ>
__device__ __forceinline__ float2 twmul(float2 a, float2 w) {
    return CONJ ? cmulc(a, w) : cmul(a, w);
}
__device__ __forceinline__ float2 cadd(float2 a, float2 b) { return make_float2(a.x + b.x, a.y + b.y); }
__device__ __forceinline__ float2 csub(float2 a, float2 b) { return make_float2(a.x - b.x, a.y - b.y); }
// a * (-i), or a * conj(-i) = a * i for the inverse
template <bool CONJ>
__device__ __forceinline__ float2 mul_mi(float2 a) {
    return CONJ ? make_float2(-a.y, a.x) : make_float2(a.y, -a.x);
}

// The three stages of a fused pass on 8 elements in registers: DIT at
// half-spans H, 2H, 4H (w1: W_2H^k, w2a/w2b: W_4H^k, W_4H^(k+H), w4[r]:
// W_8H^(k+rH)) and DIF in the reverse order.
template <bool CONJ>
__device__ __forceinline__ void dit8_regs(float2 (&x)[8], float2 w1, float2 w2a, float2 w2b,
                                          const float2 (&w4)[4]) {
#pragma unroll
    for (int r = 0; r < 8; r += 2) {  // half-span H
        const float2 t = twmul<CONJ>(x[r + 1], w1);
        x[r + 1] = csub(x[r], t);
        x[r] = cadd(x[r], t);
    }
#pragma unroll
    for (int r = 0; r < 8; r += 4) {  // half-span 2H
        const float2 t0 = twmul<CONJ>(x[r + 2], w2a);
        x[r + 2] = csub(x[r], t0);
        x[r] = cadd(x[r], t0);
        const float2 t1 = twmul<CONJ>(x[r + 3], w2b);
        x[r + 3] = csub(x[r + 1], t1);
        x[r + 1] = cadd(x[r + 1], t1);
    }
#pragma unroll
    for (int r = 0; r < 4; ++r) {  // half-span 4H
        const float2 t = twmul<CONJ>(x[r + 4], w4[r]);
        x[r + 4] = csub(x[r], t);
        x[r] = cadd(x[r], t);
    }
}
template <bool CONJ>
__device__ __forceinline__ void dif8_regs(float2 (&x)[8], float2 w1, float2 w2a, float2 w2b,
                                          const float2 (&w4)[4]) {
#pragma unroll
    for (int r = 0; r < 4; ++r) {  // half-span 4H
        const float2 d = csub(x[r], x[r + 4]);
        x[r] = cadd(x[r], x[r + 4]);
        x[r + 4] = twmul<CONJ>(d, w4[r]);
    }
#pragma unroll
    for (int r = 0; r < 8; r += 4) {  // half-span 2H
        const float2 d0 = csub(x[r], x[r + 2]);
        x[r] = cadd(x[r], x[r + 2]);
        x[r + 2] = twmul<CONJ>(d0, w2a);
        const float2 d1 = csub(x[r + 1], x[r + 3]);
        x[r + 1] = cadd(x[r + 1], x[r + 3]);
        x[r + 3] = twmul<CONJ>(d1, w2b);
    }
#pragma unroll
    for (int r = 0; r < 8; r += 2) {  // half-span H
        const float2 d = csub(x[r], x[r + 1]);
        x[r] = cadd(x[r], x[r + 1]);
        x[r + 1] = twmul<CONJ>(d, w1);
    }
}
// Half-spans 1 and 2 on a group of 4 (twiddles W_2^0 = 1, W_4^0 = 1,
// W_4^1 = -i): DIT increasing, DIF decreasing.
template <bool CONJ>
__device__ __forceinline__ void dit4_regs(float2 (&x)[4]) {
#pragma unroll
    for (int r = 0; r < 4; r += 2) {
        const float2 t = x[r + 1];
        x[r + 1] = csub(x[r], t);
        x[r] = cadd(x[r], t);
    }
    const float2 t0 = x[2];
    x[2] = csub(x[0], t0);
    x[0] = cadd(x[0], t0);
    const float2 t1 = mul_mi<CONJ>(x[3]);
    x[3] = csub(x[1], t1);
    x[1] = cadd(x[1], t1);
}
template <bool CONJ>
__device__ __forceinline__ void dif4_regs(float2 (&x)[4]) {
    const float2 d0 = csub(x[0], x[2]);
    x[0] = cadd(x[0], x[2]);
    x[2] = d0;
    const float2 d1 = csub(x[1], x[3]);
    x[1] = cadd(x[1], x[3]);
    x[3] = mul_mi<CONJ>(d1);
#pragma unroll
    for (int r = 0; r < 4; r += 2) {
        const float2 d = csub(x[r], x[r + 1]);
        x[r] = cadd(x[r], x[r + 1]);
        x[r + 1] = d;
    }
}

// One fused pass of 8-element groups base + r*H over both transforms.
template <bool DIT, bool CONJ, int N, int H, class TW>
__device__ __forceinline__ void fft_pass8(float2* X0, float2* X1, const TW& T, int tid) {
    for (int j = tid; j < N / 8; j += 256) {
        const int g = j / H, k = j - g * H;
        const int base = g * 8 * H + k;
        const float2 w1 = T(H - 1 + k);
        const float2 w2a = T(2 * H - 1 + k), w2b = T(3 * H - 1 + k);
        float2 w4[4];
#pragma unroll
        for (int r = 0; r < 4; ++r) w4[r] = T(4 * H - 1 + k + r * H);
#pragma unroll
        for (int f = 0; f < 2; ++f) {
            float2* X = f ? X1 : X0;
            float2 x[8];
#pragma unroll
            for (int r = 0; r < 8; ++r) x[r] = X[swz(base + r * H)];
            if constexpr (DIT) dit8_regs<CONJ>(x, w1, w2a, w2b, w4);
            else dif8_regs<CONJ>(x, w1, w2a, w2b, w4);
#pragma unroll
            for (int r = 0; r < 8; ++r) X[swz(base + r * H)] = x[r];
        }
    }
    __syncthreads();
}

// The middle of the filter in one LDS round trip: the forward transform's
// innermost stages (half-spans < G, on groups of G consecutive elements),
// the product with the filter (element q holds frequency brev(q)), and the
// inverse transform's first stages on the same groups.
template <int LOGN, int G, class TW>
__device__ __forceinline__ void fft_middle(float2* X0, float2* X1, const TW& T,
                                           g_cfloat* __restrict__ H, int tid) {
    constexpr int N = 1 << LOGN;
    float2 w1 = make_float2(1.0f, 0.0f), w2a = w1, w2b = w1, w4[4] = {w1, w1, w1, w1};
    if constexpr (G == 8) {  // half-span 1 twiddles of a fused radix-8 pass
        w1 = T(0);
        w2a = T(1);
        w2b = T(2);
#pragma unroll
        for (int r = 0; r < 4; ++r) w4[r] = T(3 + r);
    }
    for (int j = tid; j < N / G; j += 256) {
        float hv[G];
#pragma unroll
        for (int r = 0; r < G; ++r) {
            const int k = (int)(__brev((unsigned)(G * j + r)) >> (32 - LOGN));
            hv[r] = H[k <= N / 2 ? k : N - k];
        }
#pragma unroll
        for (int f = 0; f < 2; ++f) {
            float2* X = f ? X1 : X0;
            float2 x[G];
#pragma unroll
            for (int r = 0; r < G; ++r) x[r] = X[swz(G * j + r)];
            if constexpr (G == 8) dif8_regs<false>(x, w1, w2a, w2b, w4);
            else if constexpr (G == 4) dif4_regs<false>(x);
            else {
                const float2 d = csub(x[0], x[1]);
                x[0] = cadd(x[0], x[1]);
                x[1] = d;
            }
#pragma unroll
            for (int r = 0; r < G; ++r) x[r] = make_float2(x[r].x * hv[r], x[r].y * hv[r]);
            if constexpr (G == 8) dit8_regs<true>(x, w1, w2a, w2b, w4);
            else if constexpr (G == 4) dit4_regs<true>(x);
            else {
                const float2 d = csub(x[0], x[1]);
                x[0] = cadd(x[0], x[1]);
                x[1] = d;
            }
#pragma unroll
            for (int r = 0; r < G; ++r) X[swz(G * j + r)] = x[r];
        }
    }
    __syncthreads();
}

// Fused DIF passes at half-spans H, H/8, ... >= LAST, and DIT passes at
// H, 8H, ... < N.
template <int N, int H, int LAST, class TW>
__device__ __forceinline__ void dif8_passes(float2* X0, float2* X1, const TW& T, int tid) {
    if constexpr (H >= LAST) {
        fft_pass8<false, false, N, H>(X0, X1, T, tid);
        dif8_passes<N, H / 8, LAST>(X0, X1, T, tid);
    }
}
template <int N, int H, class TW>
__device__ __forceinline__ void dit8_passes(float2* X0, float2* X1, const TW& T, int tid) {
    if constexpr (H < N) {
        fft_pass8<true, true, N, H>(X0, X1, T, tid);
        dit8_passes<N, H * 8>(X0, X1, T, tid);
    }
}

// One (angle, z-block) work item: 4 rows -> 2 complex FFTs of n = 2^LOGN
// points in LDS.
template <int LOGN>
__device__ __forceinline__ void fft_filter_item(const FiltArgs& p, int a, int zb,
                                                float2* __restrict__ sm) {
    constexpr int n = 1 << LOGN, logn = LOGN;
    float2* X0 = sm;
    float2* X1 = sm + n;
    g_cfloat2* __restrict__ Tg = (g_cfloat2*)p.tw;  // stage-ordered, L1-resident
    auto T = [Tg](int k) {
        const f2v t = Tg[k];
        return make_float2(t.x, t.y);
    };
    g_cfloat* __restrict__ H = (g_cfloat*)p.H;
    const int tid = threadIdx.x;
    g_cfloat* rows[kZB];
#pragma unroll
    for (int r = 0; r < kZB; ++r) {
        const int z = zb * kZB + r;
        rows[r] = (z < p.nz) ? (g_cfloat*)p.sino + ((size_t)z * p.na + a) * p.bins : nullptr;
    }
    for (int j = tid; j < n; j += 256) {
        const int jj = j - p.pad;
        float v[kZB];
#pragma unroll
        for (int r = 0; r < kZB; ++r)
            v[r] = (rows[r] && jj >= 0 && jj < p.bins) ? rows[r][jj] : 0.0f;
        X0[swz(j)] = make_float2(v[0], v[1]);
        X1[swz(j)] = make_float2(v[2], v[3]);
    }
    __syncthreads();
    // Forward: decimation in frequency from natural-order input to
    // bit-reversed output; the filter is applied in that order; the inverse
    // is a decimation in time from bit-reversed input to natural-order output.
    // No element is ever moved to its bit-reversed place. Radix-2 stages are
    // fused three at a time (a thread takes the 8 elements base + r*h through
    // the stages of half-span 4h, 2h, h, or h, 2h, 4h, in registers: the same
    // operations as the separate stages with one LDS round trip).
    // The innermost 3, 1 or 2 stages of the forward transform, the product
    // with the real, even filter (H[n-k] = H[k]) and the same stages of the
    // inverse run as one pass on groups of G = 8, 2 or 4 consecutive
    // elements; the other stages run as fused radix-8 passes either side.
    constexpr int rem = logn % 3;
    constexpr int G = rem == 0 ? 8 : (1 << rem);
    dif8_passes<n, n / 8, G>(X0, X1, T, tid);
    fft_middle<logn, G>(X0, X1, T, H, tid);
    dit8_passes<n, G>(X0, X1, T, tid);
    const float inv_n = 1.0f / (float)n;
    g_float4* dst = (g_float4*)p.filt + ((size_t)zb * p.na + a) * p.sq;
    for (int i = tid; i < p.sq; i += 256) {
        const float2 u = X0[swz(i)], v = X1[swz(i)];
        const f4v o = {u.x * inv_n, u.y * inv_n, v.x * inv_n, v.y * inv_n};
        dst[i] = o;
    }
    __syncthreads();  // LDS is reused by the next item
}

template <int LOGN>
__global__ __launch_bounds__(256) void row_filter_fft_kernel(FiltArgs p) {
    extern __shared__ float2 sm[];
    fft_filter_item<LOGN>(p, blockIdx.x, blockIdx.y, sm);
}

// Persistent form: items (job, angle, z-block) of many volumes with one FFT
// length 2^LOGN, taken from an atomic counter until the list is exhausted
// (every block then returns).
template <int LOGN>
__global__ __launch_bounds__(256) void row_filter_persistent_kernel(const FiltJob* __restrict__ jobs,
                                                                    const int* __restrict__ hdr,
                                                                    int* __restrict__ counter) {
    extern __shared__ float2 sm[];
    __shared__ int item_s;
    // hdr = {first job of this launch's list, job count, item count}: read
    // from device memory so a captured launch stays valid for any batch
    jobs += hdr[0];
    const int njobs = hdr[1], total = hdr[2];
    for (;;) {
        if (threadIdx.x == 0) item_s = atomicAdd(counter, 1);
        __syncthreads();
        const int item = item_s;
        __syncthreads();
        if (item >= total) break;
        int j = 0;
        while (j + 1 < njobs && jobs[j + 1].item_begin <= item) ++j;
        const FiltJob& jb = jobs[j];
        const int local = item - jb.item_begin;
        const int zb = local / jb.a.na, a = local - zb * jb.a.na;
        fft_filter_item<LOGN>(jb.a, a, zb, sm);
    }
}

size_t fft_lds_bytes(int nfft) { return sizeof(float2) * (size_t)(2 * nfft); }

static hipError_t set_lds_attr(const void* fn) {
    return hipFuncSetAttribute(fn, hipFuncAttributeMaxDynamicSharedMemorySize,
                               (int)fft_lds_bytes(kMaxFFT));
}

template <int LOGN>
static int persistent_grid_n(int device) {
    int per_cu = 0, cus = 0;
    if (set_lds_attr((const void*)row_filter_persistent_kernel<LOGN>) != hipSuccess) return 0;
    if (hipOccupancyMaxActiveBlocksPerMultiprocessor(
            &per_cu, (const void*)row_filter_persistent_kernel<LOGN>, 256,
            fft_lds_bytes(1 << LOGN)) != hipSuccess)
        return 0;
    if (hipDeviceGetAttribute(&cus, hipDeviceAttributeMultiprocessorCount, device) != hipSuccess)
        return 0;
    return (per_cu < 1 ? 1 : per_cu) * cus;
}

template <int LOGN>
static hipError_t launch_persistent_n(const FiltJob* d_jobs, const int* d_hdr, int* d_counter,
                                      int grid, hipStream_t s) {
    static std::atomic<bool> attr{false};  // per process: the attribute is set once per kernel
    if (!attr.load(std::memory_order_acquire)) {
        const hipError_t e = set_lds_attr((const void*)row_filter_persistent_kernel<LOGN>);
        if (e != hipSuccess) return e;
        attr.store(true, std::memory_order_release);
    }
    hipLaunchKernelGGL(row_filter_persistent_kernel<LOGN>, dim3(grid), dim3(256),
                       fft_lds_bytes(1 << LOGN), s, d_jobs, d_hdr, d_counter);
    return hipGetLastError();
}

static int log2_of(int n) {
    int l = 0;
    while ((1 << l) < n) ++l;
    return l;
}

// cls: the transform class of every job in the launch's list: an FFT length
// (64 .. kMaxFFT) of the LDS form, or -R of the wave form.
int filter_persistent_grid(int cls, int device) {
    if (cls < 0) return filter_wave_persistent_grid(-cls, device);
    switch (log2_of(cls)) {
        case 6: return persistent_grid_n<6>(device);
        case 7: return persistent_grid_n<7>(device);
        case 8: return persistent_grid_n<8>(device);
        case 9: return persistent_grid_n<9>(device);
        case 10: return persistent_grid_n<10>(device);
        case 11: return persistent_grid_n<11>(device);
        case 12: return persistent_grid_n<12>(device);
        default: return 0;
    }
}

hipError_t launch_row_filter_persistent(const FiltJob* d_jobs, const int* d_hdr, int* d_counter,
                                        int cls, int grid, hipStream_t s) {
    if (cls < 0) return launch_row_filter_wave_persistent(d_jobs, d_hdr, d_counter, -cls, grid, s);
    switch (log2_of(cls)) {
        case 6: return launch_persistent_n<6>(d_jobs, d_hdr, d_counter, grid, s);
        case 7: return launch_persistent_n<7>(d_jobs, d_hdr, d_counter, grid, s);
        case 8: return launch_persistent_n<8>(d_jobs, d_hdr, d_counter, grid, s);
        case 9: return launch_persistent_n<9>(d_jobs, d_hdr, d_counter, grid, s);
        case 10: return launch_persistent_n<10>(d_jobs, d_hdr, d_counter, grid, s);
        case 11: return launch_persistent_n<11>(d_jobs, d_hdr, d_counter, grid, s);
        case 12: return launch_persistent_n<12>(d_jobs, d_hdr, d_counter, grid, s);
        default: return hipErrorInvalidValue;
    }
}

hipError_t launch_twiddles(int n, float2* d_tw, hipStream_t s) {
    const int cnt = n - 1;
    hipLaunchKernelGGL(twiddle_kernel, dim3((cnt + 127) / 128), dim3(128), 0, s, n, d_tw);
    return hipGetLastError();
}

FiltArgs make_filt_args(const float* d_sino, int nz, int na, int bins, int sq, int pad, int nfft,
                        const float* d_H, const float2* d_tw, float4* d_filt) {
    FiltArgs p;
    p.sino = d_sino;
    p.nz = nz;
    p.na = na;
    p.bins = bins;
    p.sq = sq;
    p.pad = pad;
    p.nfft = nfft;
    p.logn = 0;
    while ((1 << p.logn) < nfft) ++p.logn;
    p.H = d_H;
    p.tw = d_tw;
    p.filt = d_filt;
    p.wave_r = 0;
    return p;
}

template <int LOGN>
static hipError_t launch_fft_n(const FiltArgs& p, int nzb, hipStream_t s) {
    static std::atomic<bool> attr_set{false};  // > 64 KB of dynamic LDS must be opted into
    if (!attr_set.load(std::memory_order_acquire)) {
        const hipError_t e = set_lds_attr((const void*)row_filter_fft_kernel<LOGN>);
        if (e != hipSuccess) return e;
        attr_set.store(true, std::memory_order_release);
    }
    hipLaunchKernelGGL(row_filter_fft_kernel<LOGN>, dim3(p.na, nzb), dim3(256),
                       fft_lds_bytes(1 << LOGN), s, p);
    return hipGetLastError();
}

hipError_t launch_row_filter_fft(const float* d_sino, int nz, int na, int bins, int sq, int pad,
                                 int nfft, const float* d_H, const float2* d_tw, float4* d_filt,
                                 hipStream_t s) {
    if (nfft < 64 || nfft > kMaxFFT || (nfft & (nfft - 1))) return hipErrorInvalidValue;
    const int nzb = (nz + kZB - 1) / kZB;
    const FiltArgs p = make_filt_args(d_sino, nz, na, bins, sq, pad, nfft, d_H, d_tw, d_filt);
    switch (p.logn) {
        case 6: return launch_fft_n<6>(p, nzb, s);
        case 7: return launch_fft_n<7>(p, nzb, s);
        case 8: return launch_fft_n<8>(p, nzb, s);
        case 9: return launch_fft_n<9>(p, nzb, s);
        case 10: return launch_fft_n<10>(p, nzb, s);
        case 11: return launch_fft_n<11>(p, nzb, s);
        default: return launch_fft_n<12>(p, nzb, s);
    }
}

}  // namespace cthip
