// filter.hip — row filter of the FBP: filter design (FP64) and the padded
// FFT convolution of every sinogram row.
//
// Reference semantics (SerialBackend.cpp:150-195, FilterDesign.cpp:13-111,
// FFT.cpp:51-120): each row of `bins` samples is placed at offset
// pad = sq/2 - bins/2 in a zero row of nfft = max(64, nextPow2(2*sq)),
// transformed, multiplied by the real filter H (n/2+1 taps, H[0] = 0),
// transformed back, and the first sq samples are kept. The reference's
// inverse never reads the Nyquist bin, so the effective filter also has
// H[n/2] = 0; the design here zeroes it.
//
// Design: FP64 on the device, by direct summation of the DFT of the
// reference's float kernel f (f[0]=0.25, f[odd m] = -1/(pi m)^2 rounded to
// float), times the window, stored as float.
//
// Filtering: the filter is real and even, so two real rows ride in one
// complex FFT (row A in the real part, row B in the imaginary part) and come
// back separated. A block owns one angle of one z-block (4 rows -> 2 complex
// FFTs) held in LDS: decimation-in-time forward from bit-reversed input,
// multiply by H in natural order, decimation-in-frequency inverse whose
// output is read back at bit-reversed positions. Twiddles are read from a
// stage-ordered table in global memory (small enough to stay in L1), at
// consecutive addresses across a wave.
// Output is written z-innermost (float4 = 4 slices per bin), the layout the
// backprojection stages from.
#include "ct_internal.h"

namespace cthip {

__global__ void filter_design_kernel(int n, int type, float* __restrict__ H) {
    const int k = blockIdx.x * blockDim.x + threadIdx.x;
    const int nh = n / 2;
    if (k > nh) return;
    const double pi = 3.14159265358979323846;
    double acc = 0.25;  // f[0] * cos(0)
    for (int j = 1; j < n; j += 2) {
        const int m = (j < nh) ? j : n - j;
        const double t = 1.0 / (pi * (double)m);
        const double fj = (double)(float)(-1.0 * t * t);
        const long long r = ((long long)j * k) % n;
        acc += fj * cos(2.0 * pi * (double)r / (double)n);
    }
    double h = 2.0 * acc;
    const double freq = (double)k / (double)n;
    switch (type) {
        case 0: break;                                                  // Ramp
        case 2: h *= 0.54 + 0.46 * cos(pi * freq / 0.5); break;        // Hamming
        case 3: h *= cos(pi * freq); break;                             // Cosine
        case 4: h *= 0.5 + 0.5 * cos(pi * freq / 0.5); break;          // Hann
        case 5: {                                                        // Bartlett
            const double r = fabs(freq) / 0.5;
            h *= 1.0 - (r < 1.0 ? r : 1.0);
        } break;
        case 1:
        default: {  // Shepp-Logan (also every unknown id)
            const double om = pi * freq;
            if (k > 0) h *= sin(om) / om;
        } break;
    }
    if (k == 0 || k == nh) h = 0.0;
    H[k] = (float)h;
}

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
        const int br = (int)(__brev((unsigned)j) >> (32 - logn));
        X0[br] = make_float2(v[0], v[1]);
        X1[br] = make_float2(v[2], v[3]);
    }
    __syncthreads();
    // Forward: decimation in time from bit-reversed input, natural-order
    // output. Radix-2 stages of half-span h are fused in pairs (h, 2h): a
    // thread takes the quadruple base + {0, h, 2h, 3h} through both stages in
    // registers, the same operations as two radix-2 stages with half the LDS
    // round trips. An odd number of stages starts with one radix-2 stage.
    int h = 1;
    if (logn & 1) {
        for (int b = tid; b < nh; b += 256) {
            const int i0 = 2 * b, i1 = i0 + 1;
            const float2 u0 = X0[i0], v0 = X0[i1], u1 = X1[i0], v1 = X1[i1];
            X0[i0] = make_float2(u0.x + v0.x, u0.y + v0.y);
            X0[i1] = make_float2(u0.x - v0.x, u0.y - v0.y);
            X1[i0] = make_float2(u1.x + v1.x, u1.y + v1.y);
            X1[i1] = make_float2(u1.x - v1.x, u1.y - v1.y);
        }
        __syncthreads();
        h = 2;
    }
    const int nq = n / 4;
    for (; h < n; h <<= 2) {
        for (int j = tid; j < nq; j += 256) {
            const int g = j / h, k = j - g * h;
            const int e0 = g * 4 * h + k, e1 = e0 + h, e2 = e0 + 2 * h, e3 = e0 + 3 * h;
            const float2 w1 = T(h - 1 + k), w2 = T(2 * h - 1 + k), w3 = T(3 * h - 1 + k);
#pragma unroll
            for (int f = 0; f < 2; ++f) {
                float2* X = f ? X1 : X0;
                const float2 x0 = X[e0], x1 = cmul(X[e1], w1), x2 = X[e2], x3 = cmul(X[e3], w1);
                const float2 a0 = make_float2(x0.x + x1.x, x0.y + x1.y);
                const float2 a1 = make_float2(x0.x - x1.x, x0.y - x1.y);
                const float2 a2 = cmul(make_float2(x2.x + x3.x, x2.y + x3.y), w2);
                const float2 a3 = cmul(make_float2(x2.x - x3.x, x2.y - x3.y), w3);
                X[e0] = make_float2(a0.x + a2.x, a0.y + a2.y);
                X[e2] = make_float2(a0.x - a2.x, a0.y - a2.y);
                X[e1] = make_float2(a1.x + a3.x, a1.y + a3.y);
                X[e3] = make_float2(a1.x - a3.x, a1.y - a3.y);
            }
        }
        __syncthreads();
    }
    // multiply by the real, even filter (H[n-k] = H[k])
    for (int k = tid; k < n; k += 256) {
        const float hv = H[k <= nh ? k : n - k];
        X0[k] = make_float2(X0[k].x * hv, X0[k].y * hv);
        X1[k] = make_float2(X1[k].x * hv, X1[k].y * hv);
    }
    __syncthreads();
    // Inverse: decimation in frequency from natural-order input, bit-reversed
    // output; stages of half-span 2h and h fused in pairs, then (odd count)
    // one final radix-2 stage of half-span 1.
    for (h = n / 4; h >= 1; h >>= 2) {
        for (int j = tid; j < nq; j += 256) {
            const int g = j / h, k = j - g * h;
            const int e0 = g * 4 * h + k, e1 = e0 + h, e2 = e0 + 2 * h, e3 = e0 + 3 * h;
            const float2 w2 = T(2 * h - 1 + k), w3 = T(3 * h - 1 + k), w1 = T(h - 1 + k);
#pragma unroll
            for (int f = 0; f < 2; ++f) {
                float2* X = f ? X1 : X0;
                const float2 x0 = X[e0], x1 = X[e1], x2 = X[e2], x3 = X[e3];
                const float2 a0 = make_float2(x0.x + x2.x, x0.y + x2.y);
                const float2 a2 = cmulc(make_float2(x0.x - x2.x, x0.y - x2.y), w2);
                const float2 a1 = make_float2(x1.x + x3.x, x1.y + x3.y);
                const float2 a3 = cmulc(make_float2(x1.x - x3.x, x1.y - x3.y), w3);
                X[e0] = make_float2(a0.x + a1.x, a0.y + a1.y);
                X[e1] = cmulc(make_float2(a0.x - a1.x, a0.y - a1.y), w1);
                X[e2] = make_float2(a2.x + a3.x, a2.y + a3.y);
                X[e3] = cmulc(make_float2(a2.x - a3.x, a2.y - a3.y), w1);
            }
        }
        __syncthreads();
        if (h == 1) break;
        if (h == 2) {  // odd number of stages: finish with half-span 1
            for (int b = tid; b < nh; b += 256) {
                const int i0 = 2 * b, i1 = i0 + 1;
                const float2 u0 = X0[i0], v0 = X0[i1], u1 = X1[i0], v1 = X1[i1];
                X0[i0] = make_float2(u0.x + v0.x, u0.y + v0.y);
                X0[i1] = make_float2(u0.x - v0.x, u0.y - v0.y);
                X1[i0] = make_float2(u1.x + v1.x, u1.y + v1.y);
                X1[i1] = make_float2(u1.x - v1.x, u1.y - v1.y);
            }
            __syncthreads();
            break;
        }
    }
    const float inv_n = 1.0f / (float)n;
    g_float4* dst = (g_float4*)p.filt + ((size_t)zb * p.na + a) * p.sq;
    for (int i = tid; i < p.sq; i += 256) {
        const int br = (int)(__brev((unsigned)i) >> (32 - logn));
        const float2 u = X0[br], v = X1[br];
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

hipError_t launch_filter_design(int n, int type, float* d_H, hipStream_t s) {
    const int cnt = n / 2 + 1;
    hipLaunchKernelGGL(filter_design_kernel, dim3((cnt + 127) / 128), dim3(128), 0, s, n, type,
                       d_H);
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
