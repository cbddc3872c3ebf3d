// metrics.hip — reconstruction error metrics of a whole volume, on the device.
//
// Reference semantics (CtReconstructionController.cpp:1199-1277): over every
// voxel, diff = recon - orig in float; MSE and MAE accumulate
// double(diff)^2 and |double(diff)| in double and are divided by the voxel
// count; the largest |diff| (float) and the minimum and maximum of the
// original (float) complete the set. The host finishes PSNR and the
// max-difference rule from these sums (capi.cpp).
//
// One pass reads each volume once: 8 bytes per voxel, an HBM-bound stream.
// Every thread walks float4s grid-stride with two pairs of loads in flight
// and keeps double sums and float extrema (1024 blocks and two pairs measured
// fastest of 1024-8192 blocks x 2/4/8 pairs: 5.57 TB/s, level with a
// read-only stream of the same two volumes); a block reduces them with wave
// shuffles and LDS into one partial; one block then adds the partials in
// block order. The grid is a function of the voxel count only, so a call
// returns the same bits every time. double(diff)^2 is exact (24-bit
// operands), so the sums differ from the reference's per-thread OpenMP sums
// only in the order of the additions.
#include "ct_internal.h"

#include <cfloat>

namespace cthip {

namespace {

struct Acc {
    double s2, s1;
    float mx, lo, hi;
};

__device__ __forceinline__ void acc_one(Acc& a, float o, float r) {
    const float d = r - o;
    const double dd = (double)d;
    a.s2 = fma(dd, dd, a.s2);  // dd*dd is exact: same value as the separate multiply
    a.s1 += fabs(dd);
    a.mx = fmaxf(a.mx, fabsf(d));  // a NaN never replaces the running value,
    a.lo = fminf(a.lo, o);         // as with std::max / std::min
    a.hi = fmaxf(a.hi, o);
}

__device__ __forceinline__ void acc_four(Acc& a, const float4& o, const float4& r) {
    acc_one(a, o.x, r.x);
    acc_one(a, o.y, r.y);
    acc_one(a, o.z, r.z);
    acc_one(a, o.w, r.w);
}

__device__ __forceinline__ void acc_merge(Acc& a, const Acc& b) {
    a.s2 += b.s2;
    a.s1 += b.s1;
    a.mx = fmaxf(a.mx, b.mx);
    a.lo = fminf(a.lo, b.lo);
    a.hi = fmaxf(a.hi, b.hi);
}

// Reduces a 256-thread block's values into thread 0's (fixed shuffle tree,
// then waves in order).
__device__ __forceinline__ void block_reduce(Acc& a) {
    for (int off = 32; off > 0; off >>= 1) {
        Acc b;
        b.s2 = __shfl_xor(a.s2, off);
        b.s1 = __shfl_xor(a.s1, off);
        b.mx = __shfl_xor(a.mx, off);
        b.lo = __shfl_xor(a.lo, off);
        b.hi = __shfl_xor(a.hi, off);
        acc_merge(a, b);
    }
    __shared__ Acc ws[4];
    const int wv = threadIdx.x >> 6;
    if ((threadIdx.x & 63) == 0) ws[wv] = a;
    __syncthreads();
    if (threadIdx.x == 0) {
        a = ws[0];
        for (int i = 1; i < 4; ++i) acc_merge(a, ws[i]);
    }
}

__device__ __forceinline__ Acc acc_init() {
    Acc a;
    a.s2 = 0.0;
    a.s1 = 0.0;
    a.mx = 0.0f;
    a.lo = FLT_MAX;
    a.hi = -FLT_MAX;
    return a;
}

}  // namespace

__global__ __launch_bounds__(256) void metrics_kernel(const float* __restrict__ orig,
                                                      const float* __restrict__ recon, size_t n,
                                                      MetricsSums* __restrict__ part) {
    Acc a = acc_init();
    const size_t n4 = n / 4;
    const float4* __restrict__ o4 = reinterpret_cast<const float4*>(orig);
    const float4* __restrict__ r4 = reinterpret_cast<const float4*>(recon);
    const size_t stride = (size_t)gridDim.x * blockDim.x;
    size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
    for (; i + stride < n4; i += 2 * stride) {
        float4 o[2], r[2];
#pragma unroll
        for (int k = 0; k < 2; ++k) {
            o[k] = o4[i + k * stride];
            r[k] = r4[i + k * stride];
        }
#pragma unroll
        for (int k = 0; k < 2; ++k) acc_four(a, o[k], r[k]);
    }
    for (; i < n4; i += stride) acc_four(a, o4[i], r4[i]);
    // the n % 4 voxels past the last float4
    if (blockIdx.x == 0 && threadIdx.x < (unsigned)(n - 4 * n4)) {
        const size_t j = 4 * n4 + threadIdx.x;
        acc_one(a, orig[j], recon[j]);
    }
    block_reduce(a);
    if (threadIdx.x == 0) part[blockIdx.x] = MetricsSums{a.s2, a.s1, a.mx, a.lo, a.hi, 0.0f};
}

__global__ __launch_bounds__(256) void metrics_final_kernel(const MetricsSums* __restrict__ part,
                                                            int nparts,
                                                            MetricsSums* __restrict__ out) {
    Acc a = acc_init();
    for (int i = threadIdx.x; i < nparts; i += 256) {
        const MetricsSums p = part[i];
        acc_merge(a, Acc{p.s2, p.s1, p.mx, p.lo, p.hi});
    }
    block_reduce(a);
    if (threadIdx.x == 0) *out = MetricsSums{a.s2, a.s1, a.mx, a.lo, a.hi, 0.0f};
}

int metrics_blocks(size_t n) {
    const size_t n4 = n / 4;
    const size_t want = (n4 + 2 * 256 - 1) / (2 * 256);
    return (int)std::max<size_t>(1, std::min<size_t>(kMetricsBlocks, want));
}

hipError_t launch_metrics(const float* d_orig, const float* d_recon, size_t n,
                          MetricsSums* d_part, MetricsSums* d_out, hipStream_t s) {
    const int nb = metrics_blocks(n);
    hipLaunchKernelGGL(metrics_kernel, dim3(nb), dim3(256), 0, s, d_orig, d_recon, n, d_part);
    hipLaunchKernelGGL(metrics_final_kernel, dim3(1), dim3(256), 0, s, d_part, nb, d_out);
    return hipGetLastError();
}

}  // namespace cthip
