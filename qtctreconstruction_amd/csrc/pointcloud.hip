// pointcloud.hip — threshold compaction of a volume into (x, y, z, hu) points.
//
// Reference semantics (SerialBackend.cpp:325-346): every voxel with
// hu > threshold, in z, y, x order, with axis coordinates taken from the
// volume's coordinate vectors when present and from the indices otherwise.
//
// Two passes over blocks of kPC consecutive voxels: count per block, an
// exclusive scan of the block counts, then an ordered write in which each
// wave ranks its lanes with a ballot + popcount. Output order equals raster
// order, so the result is identical to the sequential loop.
#include "ct_internal.h"

namespace cthip {

constexpr int kPC = 4096;  // voxels per block (256 threads x 16)

__global__ __launch_bounds__(256) void pc_count_kernel(const float* __restrict__ v, size_t n,
                                                       float thr, unsigned* __restrict__ counts) {
    const size_t base = (size_t)blockIdx.x * kPC;
    unsigned cnt = 0;
#pragma unroll 4
    for (int k = 0; k < kPC / 256; ++k) {
        const size_t i = base + (size_t)k * 256 + threadIdx.x;
        cnt += (i < n && v[i] > thr) ? 1u : 0u;
    }
    for (int off = 32; off > 0; off >>= 1) cnt += (unsigned)__shfl_xor((int)cnt, off);
    __shared__ unsigned ws[4];
    if ((threadIdx.x & 63) == 0) ws[threadIdx.x >> 6] = cnt;
    __syncthreads();
    if (threadIdx.x == 0) counts[blockIdx.x] = ws[0] + ws[1] + ws[2] + ws[3];
}

__global__ __launch_bounds__(256) void pc_write_kernel(const float* __restrict__ v, size_t n,
                                                       int w, int h, float thr,
                                                       const unsigned* __restrict__ offs,
                                                       const float* __restrict__ xc,
                                                       const float* __restrict__ yc,
                                                       const float* __restrict__ zc,
                                                       float4* __restrict__ out) {
    __shared__ unsigned ws[4];
    const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
    const size_t base = (size_t)blockIdx.x * kPC;
    unsigned run = offs[blockIdx.x];
    const size_t plane = (size_t)w * h;
    for (int k = 0; k < kPC / 256; ++k) {
        const size_t i = base + (size_t)k * 256 + threadIdx.x;
        const float val = (i < n) ? v[i] : 0.0f;
        const bool pred = (i < n) && val > thr;
        const unsigned long long bal = __ballot(pred);
        const unsigned below = (unsigned)__popcll(bal & ((1ull << lane) - 1ull));
        if (lane == 0) ws[wave] = (unsigned)__popcll(bal);
        __syncthreads();
        unsigned woff = 0, total = 0;
        for (int q = 0; q < 4; ++q) {
            woff += (q < wave) ? ws[q] : 0u;
            total += ws[q];
        }
        if (pred) {
            const int z = (int)(i / plane);
            const size_t r = i - (size_t)z * plane;
            const int y = (int)(r / w), x = (int)(r - (size_t)y * w);
            out[run + woff + below] = make_float4(xc ? xc[x] : (float)x, yc ? yc[y] : (float)y,
                                                  zc ? zc[z] : (float)z, val);
        }
        run += total;
        __syncthreads();
    }
}

unsigned pc_blocks(size_t n) { return (unsigned)((n + kPC - 1) / kPC); }

hipError_t launch_pc_count(const float* d_v, size_t n, float thr, unsigned* d_counts,
                           hipStream_t s) {
    hipLaunchKernelGGL(pc_count_kernel, dim3(pc_blocks(n)), dim3(256), 0, s, d_v, n, thr, d_counts);
    return hipGetLastError();
}

hipError_t launch_pc_write(const float* d_v, size_t n, int w, int h, float thr,
                           const unsigned* d_offs, const float* d_xc, const float* d_yc,
                           const float* d_zc, float4* d_out, hipStream_t s) {
    hipLaunchKernelGGL(pc_write_kernel, dim3(pc_blocks(n)), dim3(256), 0, s, d_v, n, w, h, thr,
                       d_offs, d_xc, d_yc, d_zc, d_out);
    return hipGetLastError();
}

}  // namespace cthip
