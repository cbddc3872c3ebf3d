// phantom.hip — seeded additive-ellipsoid phantom on gfx950.
//
// Replaces the reference's per-slice generator (Generator3DGPU.cu:80-115,
// one launch plus a synchronous D2H per slice) with one launch over the whole
// volume, written straight into device memory. Voxel values are bit-equal to
// the host restatement: same float operations in the same order, contraction
// off (the library is built with -ffp-contract=off).
//
// Each block covers a 64x4 (x,y) tile of one z slice. Ellipsoids whose
// bounding box misses the tile are skipped with a wave-uniform test; skipping
// is exact because a voxel outside the box cannot pass the membership test
// (the box carries a 1e-3 margin, far above the float error of the form).
#include "ct_internal.h"

namespace cthip {

__device__ __forceinline__ float lin_coord(int i, int n, float step) {
    // utils::linspace (Utils.cpp:12-26): start + step * i, start = -1
    return (n == 1) ? -1.0f : -1.0f + step * (float)i;
}

__global__ __launch_bounds__(256) void phantom_kernel(const PhEll* __restrict__ e, int n, int nx,
                                                      int ny, int nz, float* __restrict__ vol) {
    const int x0 = blockIdx.x * 64, y0 = blockIdx.y * 4;
    const int x = x0 + (threadIdx.x & 63);
    const int y = y0 + (threadIdx.x >> 6);
    const int z = blockIdx.z;
    const float stx = 2.0f / (float)(nx - 1);
    const float sty = 2.0f / (float)(ny - 1);
    const float stz = 2.0f / (float)(nz - 1);
    const float xf = lin_coord(x, nx, stx);
    const float yf = lin_coord(y, ny, sty);
    const float zf = lin_coord(z, nz, stz);
    // tile coordinate box (linspace is increasing)
    const float tx0 = lin_coord(x0, nx, stx);
    const float tx1 = lin_coord(min(x0 + 63, nx - 1), nx, stx);
    const float ty0 = lin_coord(y0, ny, sty);
    const float ty1 = lin_coord(min(y0 + 3, ny - 1), ny, sty);
    float v = 0.0f;
    for (int k = 0; k < n; ++k) {
        const PhEll& q = e[k];
        if (tx1 < q.lo[0] || tx0 > q.hi[0] || ty1 < q.lo[1] || ty0 > q.hi[1] || zf < q.lo[2] ||
            zf > q.hi[2])
            continue;
        const float dx = xf - q.cx;
        const float dy = yf - q.cy;
        const float dz = zf - q.cz;
        const float q0 = q.m[0] * dx + q.m[1] * dy + q.m[2] * dz;
        const float q1 = q.m[3] * dx + q.m[4] * dy + q.m[5] * dz;
        const float q2 = q.m[6] * dx + q.m[7] * dy + q.m[8] * dz;
        const float f = q0 * q0 * q.ia[0] + q1 * q1 * q.ia[1] + q2 * q2 * q.ia[2];
        if (f <= 1.0f) v += q.rho;
    }
    if (x < nx && y < ny) vol[((size_t)z * ny + y) * nx + x] = v;
}

// Synthetic acquisition of a sinogram p (in place): counts ~ Poisson(I0 e^{-kappa p}),
// then p~ = -ln(max(counts,1)/I0)/kappa. The Poisson draw uses its normal
// approximation round(lam + sqrt(lam) z), z from Box-Muller on a counter-based
// hash of (seed, index); with kappa chosen so that kappa*p <= ~2 the mean is
// >= I0 e^-2 (1353 at I0 = 1e4), where the approximation is tight.
__device__ __forceinline__ unsigned long long mix64(unsigned long long x) {
    x += 0x9E3779B97F4A7C15ull;
    x = (x ^ (x >> 30)) * 0xBF58476D1CE4E5B9ull;
    x = (x ^ (x >> 27)) * 0x94D049BB133111EBull;
    return x ^ (x >> 31);
}

__global__ void noisy_log_kernel(float* __restrict__ p, size_t n, float I0, float kappa,
                                 unsigned long long seed) {
    const size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    const unsigned long long h = mix64(seed ^ mix64((unsigned long long)i));
    const float u1 = ((float)(h >> 40) + 0.5f) * (1.0f / 16777216.0f);
    const float u2 = (float)((h >> 16) & 0xffffff) * (1.0f / 16777216.0f);
    const float z = sqrtf(-2.0f * logf(u1)) * cospif(2.0f * u2);
    const float lam = I0 * expf(-kappa * p[i]);
    const float cnt = fmaxf(rintf(lam + sqrtf(lam) * z), 1.0f);
    p[i] = -logf(cnt / I0) / kappa;
}

hipError_t launch_noisy_log(float* d_p, size_t n, float I0, float kappa, unsigned long long seed,
                            hipStream_t s) {
    hipLaunchKernelGGL(noisy_log_kernel, dim3((unsigned)((n + 255) / 256)), dim3(256), 0, s, d_p, n,
                       I0, kappa, seed);
    return hipGetLastError();
}

hipError_t launch_phantom(const PhEll* d_ell, int n, int nx, int ny, int nz, float* d_vol,
                          hipStream_t s) {
    dim3 grid((nx + 63) / 64, (ny + 3) / 4, nz);
    hipLaunchKernelGGL(phantom_kernel, grid, dim3(256), 0, s, d_ell, n, nx, ny, nz, d_vol);
    return hipGetLastError();
}

}  // namespace cthip
