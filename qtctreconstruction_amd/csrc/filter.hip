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

__device__ __forceinline__ float2 cmul(float2 a, float2 b) {
    return make_float2(a.x * b.x - a.y * b.y, a.x * b.y + a.y * b.x);
}
__device__ __forceinline__ float2 cmulc(float2 a, float2 b) {  // a * conj(b)
    return make_float2(a.x * b.x + a.y * b.y, a.y * b.x - a.x * b.y);
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

// Radix-2 butterfly stages on the two LDS transforms X0, X1 of n points.
// Stage of half-span m, element e at position q = e mod 2m < m paired with
// e + m; W = W_2m^q (stage order: W_M^j at M/2 - 1 + j), conjugated when
// CONJ (the inverse transform):
//   DIT (increasing spans, bit-reversed in, natural out):
//       t = W*b;  a' = a + t;  b' = a - t
//   DIF (decreasing spans, natural in, bit-reversed out):
//       a' = a + b;  b' = W*(a - b)
// A fused pass applies 1, 2 or 3 consecutive stages to groups of 2, 4 or 8
// elements in registers, then writes them back; every stage's operations
// and their order are those of the single stage.
template <bool CONJ>
__device__ __forceinline__ float2 twmul(float2 a, float2 w) {
    return CONJ ? cmulc(a, w) : cmul(a, w);
}

template <bool CONJ, class TW>
__device__ __forceinline__ void fft_dit8(float2* X0, float2* X1, int n, int h, const TW& T,
                                         int tid) {
    for (int j = tid; j < n / 8; j += 256) {
        const int g = j / h, k = j - g * h;
        const int base = g * 8 * h + k;
        const float2 w1 = T(h - 1 + k);
        const float2 w2a = T(2 * h - 1 + k), w2b = T(3 * h - 1 + k);
        float2 w4[4];
#pragma unroll
        for (int r = 0; r < 4; ++r) w4[r] = T(4 * h - 1 + k + r * h);
#pragma unroll
        for (int f = 0; f < 2; ++f) {
            float2* X = f ? X1 : X0;
            float2 x[8];
#pragma unroll
            for (int r = 0; r < 8; ++r) x[r] = X[swz(base + r * h)];
#pragma unroll
            for (int r = 0; r < 8; r += 2) {  // half-span h
                const float2 t = twmul<CONJ>(x[r + 1], w1);
                x[r + 1] = make_float2(x[r].x - t.x, x[r].y - t.y);
                x[r] = make_float2(x[r].x + t.x, x[r].y + t.y);
            }
#pragma unroll
            for (int r = 0; r < 8; r += 4) {  // half-span 2h
                const float2 t0 = twmul<CONJ>(x[r + 2], w2a);
                x[r + 2] = make_float2(x[r].x - t0.x, x[r].y - t0.y);
                x[r] = make_float2(x[r].x + t0.x, x[r].y + t0.y);
                const float2 t1 = twmul<CONJ>(x[r + 3], w2b);
                x[r + 3] = make_float2(x[r + 1].x - t1.x, x[r + 1].y - t1.y);
                x[r + 1] = make_float2(x[r + 1].x + t1.x, x[r + 1].y + t1.y);
            }
#pragma unroll
            for (int r = 0; r < 4; ++r) {  // half-span 4h
                const float2 t = twmul<CONJ>(x[r + 4], w4[r]);
                x[r + 4] = make_float2(x[r].x - t.x, x[r].y - t.y);
                x[r] = make_float2(x[r].x + t.x, x[r].y + t.y);
            }
#pragma unroll
            for (int r = 0; r < 8; ++r) X[swz(base + r * h)] = x[r];
        }
    }
    __syncthreads();
}

template <bool CONJ, class TW>
__device__ __forceinline__ void fft_dif8(float2* X0, float2* X1, int n, int h, const TW& T,
                                         int tid) {
    for (int j = tid; j < n / 8; j += 256) {
        const int g = j / h, k = j - g * h;
        const int base = g * 8 * h + k;
        const float2 w1 = T(h - 1 + k);
        const float2 w2a = T(2 * h - 1 + k), w2b = T(3 * h - 1 + k);
        float2 w4[4];
#pragma unroll
        for (int r = 0; r < 4; ++r) w4[r] = T(4 * h - 1 + k + r * h);
#pragma unroll
        for (int f = 0; f < 2; ++f) {
            float2* X = f ? X1 : X0;
            float2 x[8];
#pragma unroll
            for (int r = 0; r < 8; ++r) x[r] = X[swz(base + r * h)];
#pragma unroll
            for (int r = 0; r < 4; ++r) {  // half-span 4h
                const float2 d = make_float2(x[r].x - x[r + 4].x, x[r].y - x[r + 4].y);
                x[r] = make_float2(x[r].x + x[r + 4].x, x[r].y + x[r + 4].y);
                x[r + 4] = twmul<CONJ>(d, w4[r]);
            }
#pragma unroll
            for (int r = 0; r < 8; r += 4) {  // half-span 2h
                const float2 d0 = make_float2(x[r].x - x[r + 2].x, x[r].y - x[r + 2].y);
                x[r] = make_float2(x[r].x + x[r + 2].x, x[r].y + x[r + 2].y);
                x[r + 2] = twmul<CONJ>(d0, w2a);
                const float2 d1 = make_float2(x[r + 1].x - x[r + 3].x, x[r + 1].y - x[r + 3].y);
                x[r + 1] = make_float2(x[r + 1].x + x[r + 3].x, x[r + 1].y + x[r + 3].y);
                x[r + 3] = twmul<CONJ>(d1, w2b);
            }
#pragma unroll
            for (int r = 0; r < 8; r += 2) {  // half-span h
                const float2 d = make_float2(x[r].x - x[r + 1].x, x[r].y - x[r + 1].y);
                x[r] = make_float2(x[r].x + x[r + 1].x, x[r].y + x[r + 1].y);
                x[r + 1] = twmul<CONJ>(d, w1);
            }
#pragma unroll
            for (int r = 0; r < 8; ++r) X[swz(base + r * h)] = x[r];
        }
    }
    __syncthreads();
}

template <bool CONJ, class TW>
__device__ __forceinline__ void fft_dit4(float2* X0, float2* X1, int n, int h, const TW& T,
                                         int tid) {
    for (int j = tid; j < n / 4; j += 256) {
        const int g = j / h, k = j - g * h;
        const int base = g * 4 * h + k;
        const float2 w1 = T(h - 1 + k), w2a = T(2 * h - 1 + k), w2b = T(3 * h - 1 + k);
#pragma unroll
        for (int f = 0; f < 2; ++f) {
            float2* X = f ? X1 : X0;
            float2 x[4];
#pragma unroll
            for (int r = 0; r < 4; ++r) x[r] = X[swz(base + r * h)];
#pragma unroll
            for (int r = 0; r < 4; r += 2) {
                const float2 t = twmul<CONJ>(x[r + 1], w1);
                x[r + 1] = make_float2(x[r].x - t.x, x[r].y - t.y);
                x[r] = make_float2(x[r].x + t.x, x[r].y + t.y);
            }
            const float2 t0 = twmul<CONJ>(x[2], w2a);
            x[2] = make_float2(x[0].x - t0.x, x[0].y - t0.y);
            x[0] = make_float2(x[0].x + t0.x, x[0].y + t0.y);
            const float2 t1 = twmul<CONJ>(x[3], w2b);
            x[3] = make_float2(x[1].x - t1.x, x[1].y - t1.y);
            x[1] = make_float2(x[1].x + t1.x, x[1].y + t1.y);
#pragma unroll
            for (int r = 0; r < 4; ++r) X[swz(base + r * h)] = x[r];
        }
    }
    __syncthreads();
}

template <bool CONJ, class TW>
__device__ __forceinline__ void fft_dif4(float2* X0, float2* X1, int n, int h, const TW& T,
                                         int tid) {
    for (int j = tid; j < n / 4; j += 256) {
        const int g = j / h, k = j - g * h;
        const int base = g * 4 * h + k;
        const float2 w1 = T(h - 1 + k), w2a = T(2 * h - 1 + k), w2b = T(3 * h - 1 + k);
#pragma unroll
        for (int f = 0; f < 2; ++f) {
            float2* X = f ? X1 : X0;
            float2 x[4];
#pragma unroll
            for (int r = 0; r < 4; ++r) x[r] = X[swz(base + r * h)];
            const float2 d0 = make_float2(x[0].x - x[2].x, x[0].y - x[2].y);
            x[0] = make_float2(x[0].x + x[2].x, x[0].y + x[2].y);
            x[2] = twmul<CONJ>(d0, w2a);
            const float2 d1 = make_float2(x[1].x - x[3].x, x[1].y - x[3].y);
            x[1] = make_float2(x[1].x + x[3].x, x[1].y + x[3].y);
            x[3] = twmul<CONJ>(d1, w2b);
#pragma unroll
            for (int r = 0; r < 4; r += 2) {
                const float2 d = make_float2(x[r].x - x[r + 1].x, x[r].y - x[r + 1].y);
                x[r] = make_float2(x[r].x + x[r + 1].x, x[r].y + x[r + 1].y);
                x[r + 1] = twmul<CONJ>(d, w1);
            }
#pragma unroll
            for (int r = 0; r < 4; ++r) X[swz(base + r * h)] = x[r];
        }
    }
    __syncthreads();
}

// half-span 1 (twiddle 1)
__device__ __forceinline__ void fft_dit2(float2* X0, float2* X1, int n, int tid) {
    for (int b = tid; b < n / 2; b += 256) {
        const int i0 = swz(2 * b), i1 = swz(2 * b + 1);
        const float2 u0 = X0[i0], v0 = X0[i1], u1 = X1[i0], v1 = X1[i1];
        X0[i0] = make_float2(u0.x + v0.x, u0.y + v0.y);
        X0[i1] = make_float2(u0.x - v0.x, u0.y - v0.y);
        X1[i0] = make_float2(u1.x + v1.x, u1.y + v1.y);
        X1[i1] = make_float2(u1.x - v1.x, u1.y - v1.y);
    }
    __syncthreads();
}
__device__ __forceinline__ void fft_dif2(float2* X0, float2* X1, int n, int tid) {
    fft_dit2(X0, X1, n, tid);  // with twiddle 1 both forms are a + b, a - b
}

// One (angle, z-block) work item: 4 rows -> 2 complex FFTs of n points in LDS.
__device__ __forceinline__ void fft_filter_item(const FiltArgs& p, int a, int zb,
                                                float2* __restrict__ sm) {
    const int n = p.nfft, logn = p.logn;
    float2* X0 = sm;
    float2* X1 = sm + n;
    g_cfloat2* __restrict__ Tg = (g_cfloat2*)p.tw;  // stage-ordered, L1-resident
    auto T = [Tg](int k) {
        const f2v t = Tg[k];
        return make_float2(t.x, t.y);
    };
    g_cfloat* __restrict__ H = (g_cfloat*)p.H;
    const int tid = threadIdx.x;
    const int nh = n / 2;
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
    // operations as the separate stages with one LDS round trip); the 1 or 2
    // remaining stages run as one radix-2 or radix-4 pass.
    const int rem = logn % 3;
    int h = n / 8;
    for (int pass = 0; pass < logn / 3; ++pass, h >>= 3)
        fft_dif8<false>(X0, X1, n, h, T, tid);
    if (rem == 1) fft_dif2(X0, X1, n, tid);
    else if (rem == 2) fft_dif4<false>(X0, X1, n, 1, T, tid);
    // multiply by the real, even filter (H[n-k] = H[k]); element p holds
    // frequency brev(p)
    for (int q = tid; q < n; q += 256) {
        const int k = (int)(__brev((unsigned)q) >> (32 - logn));
        const float hv = H[k <= nh ? k : n - k];
        const int qs = swz(q);
        X0[qs] = make_float2(X0[qs].x * hv, X0[qs].y * hv);
        X1[qs] = make_float2(X1[qs].x * hv, X1[qs].y * hv);
    }
    __syncthreads();
    h = 1;
    if (rem == 1) {
        fft_dit2(X0, X1, n, tid);
        h = 2;
    } else if (rem == 2) {
        fft_dit4<true>(X0, X1, n, 1, T, tid);
        h = 4;
    }
    for (; h < n; h <<= 3) fft_dit8<true>(X0, X1, n, h, T, tid);
    const float inv_n = 1.0f / (float)n;
    g_float4* dst = (g_float4*)p.filt + ((size_t)zb * p.na + a) * p.sq;
    for (int i = tid; i < p.sq; i += 256) {
        const float2 u = X0[swz(i)], v = X1[swz(i)];
        const f4v o = {u.x * inv_n, u.y * inv_n, v.x * inv_n, v.y * inv_n};
        dst[i] = o;
    }
    __syncthreads();  // LDS is reused by the next item
}

__global__ __launch_bounds__(256) void row_filter_fft_kernel(FiltArgs p) {
    extern __shared__ float2 sm[];
    fft_filter_item(p, blockIdx.x, blockIdx.y, sm);
}

// Persistent form: items (job, angle, z-block) of many volumes, taken from an
// atomic counter until the list is exhausted (every block then returns).
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
        fft_filter_item(jb.a, a, zb, sm);
    }
}

size_t fft_lds_bytes(int nfft) { return sizeof(float2) * (size_t)(2 * nfft); }

static hipError_t set_lds_attr(const void* fn) {
    return hipFuncSetAttribute(fn, hipFuncAttributeMaxDynamicSharedMemorySize,
                               (int)fft_lds_bytes(kMaxFFT));
}

int filter_persistent_grid(int max_nfft, int device) {
    int per_cu = 0, cus = 0;
    if (set_lds_attr((const void*)row_filter_persistent_kernel) != hipSuccess) return 0;
    if (hipOccupancyMaxActiveBlocksPerMultiprocessor(&per_cu, (const void*)row_filter_persistent_kernel,
                                                     256, fft_lds_bytes(max_nfft)) != hipSuccess)
        return 0;
    if (hipDeviceGetAttribute(&cus, hipDeviceAttributeMultiprocessorCount, device) != hipSuccess)
        return 0;
    return (per_cu < 1 ? 1 : per_cu) * cus;
}

hipError_t launch_row_filter_persistent(const FiltJob* d_jobs, const int* d_hdr, int* d_counter,
                                        int max_nfft, int grid, hipStream_t s) {
    static bool attr = false;
    if (!attr) {
        const hipError_t e = set_lds_attr((const void*)row_filter_persistent_kernel);
        if (e != hipSuccess) return e;
        attr = true;
    }
    hipLaunchKernelGGL(row_filter_persistent_kernel, dim3(grid), dim3(256), fft_lds_bytes(max_nfft),
                       s, d_jobs, d_hdr, d_counter);
    return hipGetLastError();
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
    return p;
}

hipError_t launch_row_filter_fft(const float* d_sino, int nz, int na, int bins, int sq, int pad,
                                 int nfft, const float* d_H, const float2* d_tw, float4* d_filt,
                                 hipStream_t s) {
    if (nfft < 4 || nfft > kMaxFFT || (nfft & (nfft - 1))) return hipErrorInvalidValue;
    static bool attr_set = false;  // > 64 KB of dynamic LDS must be opted into
    if (!attr_set) {
        const hipError_t e = set_lds_attr((const void*)row_filter_fft_kernel);
        if (e != hipSuccess) return e;
        attr_set = true;
    }
    const int nzb = (nz + kZB - 1) / kZB;
    const FiltArgs p = make_filt_args(d_sino, nz, na, bins, sq, pad, nfft, d_H, d_tw, d_filt);
    hipLaunchKernelGGL(row_filter_fft_kernel, dim3(na, nzb), dim3(256), fft_lds_bytes(nfft), s, p);
    return hipGetLastError();
}

}  // namespace cthip
