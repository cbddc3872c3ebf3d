// radon.hip — per-slice normalisation and the ray-driven forward projection.
//
// The reference's oracle (SerialBackend::computeSinogram, SerialBackend.cpp:59-136)
// is a pixel-driven scatter: for every angle it walks the slice in raster order
// (y outer, x inner), skips pixels whose raw value is exactly 0, normalises
// v = (raw - min) * (1/span), and deposits v*(1-frac) / v*frac into bins
// i0 = floor(s), i0+1 (full v when i0 == bins-1, nothing when i0 < 0).
//
// Here every lane owns one detector bin of one angle (a ray) and gathers: it
// visits, in the same raster order, exactly the pixels whose floor(s) is its
// bin or the bin below, recomputing s with Serial's float expression. Each
// bin therefore receives the same float additions in the same order as in the
// scatter, so the sinogram is bit-identical to Serial's, with no atomics.
// The candidate pixels come from a conservative interval in s (+-1/32 of a
// bin); the exact floor(s) test decides.
//
// Eight slices share one lane: the geometry is slice-independent, so the index
// work is done once and the eight values arrive as two float4 loads. Rays of
// angles nearer 0/180 deg walk down columns of the row-major copy; rays of
// angles nearer 90 deg walk along rows, and read the transposed copy, so that
// adjacent lanes (adjacent bins) read adjacent addresses in both cases.
#include "ct_internal.h"

namespace cthip {

__device__ __forceinline__ unsigned f2key(float f) {
    const unsigned u = __float_as_uint(f);
    return (u & 0x80000000u) ? ~u : (u | 0x80000000u);
}
__device__ __forceinline__ float key2f(unsigned k) {
    return __uint_as_float((k & 0x80000000u) ? (k & 0x7fffffffu) : ~k);
}

__global__ void minmax_init_kernel(unsigned* keys, int nz) {
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i < nz) {
        keys[2 * i] = 0xffffffffu;
        keys[2 * i + 1] = 0u;
    }
}

// grid (chunks, nz): block-reduce a chunk of one slice, then one atomic each.
__global__ __launch_bounds__(256) void minmax_kernel(const float* __restrict__ vol,
                                                     size_t slice_elems, unsigned* keys) {
    const int z = blockIdx.y;
    const float* p = vol + (size_t)z * slice_elems;
    unsigned kmin = 0xffffffffu, kmax = 0u;
    const size_t stride = (size_t)gridDim.x * blockDim.x;
    for (size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x; i < slice_elems; i += stride) {
        const unsigned k = f2key(p[i]);
        kmin = min(kmin, k);
        kmax = max(kmax, k);
    }
    for (int off = 32; off > 0; off >>= 1) {
        kmin = min(kmin, (unsigned)__shfl_xor((int)kmin, off));
        kmax = max(kmax, (unsigned)__shfl_xor((int)kmax, off));
    }
    __shared__ unsigned smin[4], smax[4];
    const int wv = threadIdx.x >> 6;
    if ((threadIdx.x & 63) == 0) {
        smin[wv] = kmin;
        smax[wv] = kmax;
    }
    __syncthreads();
    if (threadIdx.x == 0) {
        for (int i = 1; i < 4; ++i) {
            kmin = min(kmin, smin[i]);
            kmax = max(kmax, smax[i]);
        }
        atomicMin(&keys[2 * z], kmin);
        atomicMax(&keys[2 * z + 1], kmax);
    }
}

__global__ void minmax_decode_kernel(const unsigned* keys, int nz, float* mm) {
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i < 2 * nz) mm[i] = key2f(keys[i]);
}

// block (32,8) handles a 32x32 pixel tile of one Radon z-block (kRZ = 8
// slices): reads the slices, normalises (SerialBackend.cpp:78-80,104,106),
// writes the row layout [zb][y][x][8] directly and the column layout
// [zb][x][y][8] through an LDS transpose. A pixel is 32 bytes (2 float4).
__global__ __launch_bounds__(256) void normalise_kernel(const float* __restrict__ vol, int w,
                                                        int h, int nz,
                                                        const float* __restrict__ mm,
                                                        float4* __restrict__ rowL,
                                                        float4* __restrict__ colL) {
    __shared__ float4 tile[32][33][2];
    const int zb = blockIdx.z;
    const int bx = blockIdx.x * 32, by = blockIdx.y * 32;
    float mn[kRZ], inv[kRZ];
    bool live[kRZ];
#pragma unroll
    for (int k = 0; k < kRZ; ++k) {
        const int z = zb * kRZ + k;
        live[k] = z < nz;
        const float lo = live[k] ? mm[2 * z] : 0.0f;
        const float hi = live[k] ? mm[2 * z + 1] : 1.0f;
        float span = hi - lo;
        if (span < 1e-6f) span = 1.0f;
        mn[k] = lo;
        inv[k] = 1.0f / span;
    }
    const size_t ss = (size_t)w * h;
    for (int r = threadIdx.y; r < 32; r += 8) {
        const int x = bx + threadIdx.x, y = by + r;
        float t[kRZ];
#pragma unroll
        for (int k = 0; k < kRZ; ++k) t[k] = 0.0f;
        if (x < w && y < h) {
#pragma unroll
            for (int k = 0; k < kRZ; ++k) {
                if (live[k]) {
                    const float raw = vol[(size_t)(zb * kRZ + k) * ss + (size_t)y * w + x];
                    t[k] = (raw == 0.0f) ? 0.0f : (raw - mn[k]) * inv[k];
                }
            }
        }
        const float4 v0 = make_float4(t[0], t[1], t[2], t[3]);
        const float4 v1 = make_float4(t[4], t[5], t[6], t[7]);
        if (x < w && y < h) {
            float4* d = rowL + (((size_t)zb * h + y) * w + x) * 2;
            d[0] = v0;
            d[1] = v1;
        }
        tile[r][threadIdx.x][0] = v0;
        tile[r][threadIdx.x][1] = v1;
    }
    __syncthreads();
    for (int r = threadIdx.y; r < 32; r += 8) {
        const int y = by + threadIdx.x, x = bx + r;
        if (x < w && y < h) {
            float4* d = colL + (((size_t)zb * w + x) * h + y) * 2;
            d[0] = tile[threadIdx.x][r][0];
            d[1] = tile[threadIdx.x][r][1];
        }
    }
}

// s-interval margin: the float s of SerialBackend.cpp:110 is within ~3e-4 of
// the exact value for slices up to 2048 px, so a candidate set built from
// [i-1-kDelta, i+1+kDelta] in exact arithmetic contains every pixel with
// floor(s_float) in {i-1, i}.
constexpr float kDelta = 1.0f / 32.0f;

__global__ __launch_bounds__(256) void radon_ray_kernel(
    const float4* __restrict__ rowL, const float4* __restrict__ colL, int w, int h, int nz,
    int na, int bins, const float* __restrict__ cos_t, const float* __restrict__ sin_t,
    float* __restrict__ sino) {
    const int i = blockIdx.x * 256 + threadIdx.x;  // bin owned by this lane
    const int a = blockIdx.y;
    const int zb = blockIdx.z;
    if (i >= bins) return;
    const float c = cos_t[a], sn = sin_t[a];
    const float cx = (float)(w - 1) * 0.5f;
    const float cy = (float)(h - 1) * 0.5f;
    const float dc = (float)(bins / 2);
    const bool col = fabsf(sn) > fabsf(c);
    const float4* __restrict__ base = (col ? colL : rowL) + (size_t)zb * w * h * 2;
    const size_t sx_stride = col ? (size_t)h * 2 : 2;  // float4 step for x+1
    const size_t sy_stride = col ? 2 : (size_t)w * 2;  // float4 step for y+1

    const float slo = (float)(i - 1) - kDelta - dc;  // s - dc lower bound
    const float shi = (float)(i + 1) + kDelta - dc;  // s - dc upper bound
    const float ac = fabsf(c), as = fabsf(sn);
    // factor rules of SerialBackend.cpp:111-120 for this bin
    const bool last = (i == bins - 1);
    const int im1 = (i >= 1) ? i - 1 : -0x7fffffff;

    // rows that can hold a candidate: (y-cy)*sn in [slo - cx|c|, shi + cx|c|]
    int y0 = 0, y1 = h - 1;
    if (as > 1e-20f) {
        const float lo = (slo - cx * ac) / as;
        const float hi = (shi + cx * ac) / as;
        float ya, yb;
        if (sn > 0.0f) { ya = cy + lo; yb = cy + hi; }
        else           { ya = cy - hi; yb = cy - lo; }
        y0 = max(0, (int)floorf(fmaxf(ya, -2.0f)) - 1);
        y1 = min(h - 1, (int)ceilf(fminf(yb, (float)h + 1.0f)) + 1);
    }

    // Per-row candidate interval in x: x - cx = ((s-dc) - (y-cy) sn) / c over
    // s - dc in [slo, shi]. For |c| >= |s| it is stepped from row to row,
    // x in [xa0 + y*dxa, xa0 + y*dxa + span] (|dxa| <= 1, |1/c| <= sqrt2, so
    // the stepping error, ~1e-3 px, is far below the kDelta/|c| margin);
    // steeper rays compute it per row from sy, whose error then scales with
    // the distance to the strip and vanishes where candidates exist.
    const bool has_c = ac > 1e-20f;
    const float inv_c = has_c ? 1.0f / c : 0.0f;
    float xa0 = 0.0f, dxa = 0.0f, span = (float)w;
    if (!col) {
        const float e0 = (slo + cy * sn) * inv_c, e1 = (shi + cy * sn) * inv_c;
        xa0 = cx + fminf(e0, e1);
        span = fabsf(e1 - e0);
        dxa = -sn * inv_c;
    }

    float acc[kRZ];
#pragma unroll
    for (int k = 0; k < kRZ; ++k) acc[k] = 0.0f;
    for (int y = y0; y <= y1; ++y) {
        const float sy = ((float)y - cy) * sn;
        int x0 = 0, x1 = w - 1;
        if (!col) {
            const float xa = xa0 + (float)y * dxa;
            x0 = max(0, (int)ceilf(fmaxf(xa - 0.01f, -1.0f)));
            x1 = min(w - 1, (int)floorf(fminf(xa + span + 0.01f, (float)w)));
        } else if (has_c) {
            const float lo = (slo - sy) * inv_c;
            const float hi = (shi - sy) * inv_c;
            x0 = max(0, (int)ceilf(fmaxf(cx + fminf(lo, hi), -1.0f)));
            x1 = min(w - 1, (int)floorf(fminf(cx + fmaxf(lo, hi), (float)w)));
        }
        const float4* p = base + (size_t)y * sy_stride + (size_t)x0 * sx_stride;
        float xmc = (float)x0 - cx;
        for (int x = x0; x <= x1; ++x, xmc += 1.0f, p += sx_stride) {
            // SerialBackend.cpp:110-120, same float operations and order
            const float s = xmc * c + sy + dc;
            const float fl = floorf(s);
            const int i0 = (int)fl;
            const float frac = s - fl;
            const float f_here = last ? 1.0f : 1.0f - frac;
            // non-candidates add +0, which leaves every partial sum unchanged
            const float fac = (i0 == i) ? f_here : ((i0 == im1) ? frac : 0.0f);
            const float4 v0 = p[0], v1 = p[1];
            acc[0] += v0.x * fac;
            acc[1] += v0.y * fac;
            acc[2] += v0.z * fac;
            acc[3] += v0.w * fac;
            acc[4] += v1.x * fac;
            acc[5] += v1.y * fac;
            acc[6] += v1.z * fac;
            acc[7] += v1.w * fac;
        }
    }
#pragma unroll
    for (int k = 0; k < kRZ; ++k) {
        const int z = zb * kRZ + k;
        if (z < nz) sino[((size_t)z * na + a) * bins + i] = acc[k];
    }
}

hipError_t launch_minmax_init(unsigned* d_keys, int nz, hipStream_t s) {
    hipLaunchKernelGGL(minmax_init_kernel, dim3((nz + 255) / 256), dim3(256), 0, s, d_keys, nz);
    return hipGetLastError();
}

hipError_t launch_minmax(const float* d_vol, size_t slice_elems, int nz, unsigned* d_keys,
                         hipStream_t s) {
    int chunks = (int)((slice_elems + 256 * 16 - 1) / (256 * 16));
    chunks = chunks < 1 ? 1 : (chunks > 64 ? 64 : chunks);
    hipLaunchKernelGGL(minmax_kernel, dim3(chunks, nz), dim3(256), 0, s, d_vol, slice_elems,
                       d_keys);
    return hipGetLastError();
}

hipError_t launch_minmax_decode(const unsigned* d_keys, int nz, float* d_mm, hipStream_t s) {
    hipLaunchKernelGGL(minmax_decode_kernel, dim3((2 * nz + 255) / 256), dim3(256), 0, s, d_keys,
                       nz, d_mm);
    return hipGetLastError();
}

hipError_t launch_normalise(const float* d_vol, int w, int h, int nz, const float* d_mm,
                            float4* d_row, float4* d_col, hipStream_t s) {
    const int nzb = (nz + kRZ - 1) / kRZ;
    dim3 grid((w + 31) / 32, (h + 31) / 32, nzb);
    hipLaunchKernelGGL(normalise_kernel, grid, dim3(32, 8), 0, s, d_vol, w, h, nz, d_mm, d_row,
                       d_col);
    return hipGetLastError();
}

hipError_t launch_radon(const float4* d_row, const float4* d_col, int w, int h, int nz, int na,
                        int bins, const float* d_cos, const float* d_sin, float* d_sino,
                        hipStream_t s) {
    const int nzb = (nz + kRZ - 1) / kRZ;
    dim3 grid((bins + 255) / 256, na, nzb);
    hipLaunchKernelGGL(radon_ray_kernel, grid, dim3(256), 0, s, d_row, d_col, w, h, nz, na, bins,
                       d_cos, d_sin, d_sino);
    return hipGetLastError();
}

}  // namespace cthip
